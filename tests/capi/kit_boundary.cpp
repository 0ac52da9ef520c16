// A C++ caller of the C-ABI that owns its device arrays, as a STORMM host program owns its
// PhaseSpaceSynthesis / ScoreCard / LineMinimization data: the arrays are hipMalloc'd here in the
// PsSynthesisWriter, ScoreCardWriter and LinMinWriter layouts, and the kit entry points evaluate
// and minimize on them.  Every result must equal the handle entry points' bit for bit, at 32 / 23
// bits in single precision, and at 40 / 40 (the reference's own minimization set-up) and 60 / 60
// bits (95-bit coordinates and forces) in double precision.  Prints one line per case and exits
// non-zero on a mismatch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/stormm_amd.h"

#define CHECK(expr)                                                                          \
  do {                                                                                       \
    const int _rc = (expr);                                                                  \
    if (_rc != STORMM_OK) {                                                                  \
      std::printf("FAILED %s: %d %s\n", #expr, _rc, stormm_last_error());                  \
      return 1;                                                                              \
    }                                                                                        \
  } while (0)
#define HIPCHECK(expr)                                                                       \
  do {                                                                                       \
    if ((expr) != hipSuccess) {                                                              \
      std::printf("FAILED %s\n", #expr);                                                     \
      return 1;                                                                              \
    }                                                                                        \
  } while (0)

namespace {

// A branched chain of n atoms built by a simple random walk: bonds along the chain, angles and
// dihedrals over consecutive atoms, zero-sum charges, combining-rule Lennard-Jones, OBC-II radii.
struct Molecules {
  std::vector<int32_t> atom_counts, lj_types, bond_counts, angle_counts, dihe_counts, combining;
  std::vector<double> charge, radii, screen, sigma, epsilon, xyz;
  std::vector<int32_t> lj_type, bonds, angles, dihedrals, improper;
  std::vector<double> bond_k, bond_r0, angle_k, angle_t, amp, phase, freq, scee, scnb;

  void add(int n, std::mt19937 &rng) {
    std::uniform_real_distribution<double> u(0.0, 1.0);
    atom_counts.push_back(n);
    lj_types.push_back(n);
    combining.push_back(1);
    double qsum = 0.0;
    std::vector<double> q(n);
    for (int i = 0; i < n; i++) {
      q[i] = 0.3 * (u(rng) - 0.5);
      qsum += q[i];
    }
    double x = 0.0, y = 0.0, z = 0.0;
    for (int i = 0; i < n; i++) {
      charge.push_back(q[i] - qsum / n);
      lj_type.push_back(i);
      sigma.push_back(2.0 + 1.8 * u(rng));
      epsilon.push_back(0.02 + 0.18 * u(rng));
      radii.push_back(1.2 + 0.8 * u(rng));
      screen.push_back(0.7 + 0.2 * u(rng));
      x += 1.0 + 0.4 * u(rng);
      y += 0.9 * (u(rng) - 0.5);
      z += 0.9 * (u(rng) - 0.5);
      xyz.push_back(x);
      xyz.push_back(y);
      xyz.push_back(z);
    }
    bond_counts.push_back(n - 1);
    for (int i = 0; i + 1 < n; i++) {
      bonds.push_back(i);
      bonds.push_back(i + 1);
      bond_k.push_back(200.0 + 200.0 * u(rng));
      bond_r0.push_back(1.0 + 0.5 * u(rng));
    }
    angle_counts.push_back(n - 2);
    for (int i = 0; i + 2 < n; i++) {
      angles.insert(angles.end(), {i, i + 1, i + 2});
      angle_k.push_back(40.0 + 40.0 * u(rng));
      angle_t.push_back((105.0 + 15.0 * u(rng)) * M_PI / 180.0);
    }
    dihe_counts.push_back(n - 3);
    for (int i = 0; i + 3 < n; i++) {
      dihedrals.insert(dihedrals.end(), {i, i + 1, i + 2, i + 3});
      amp.push_back(2.0 * u(rng));
      phase.push_back((i % 2) ? M_PI : 0.0);
      freq.push_back(1 + (i % 3));
      improper.push_back(0);
      scee.push_back(1.2);
      scnb.push_back(2.0);
    }
  }

  stormm_topology_batch batch() const {
    stormm_topology_batch b;
    std::memset(&b, 0, sizeof(b));
    b.system_count = static_cast<int32_t>(atom_counts.size());
    b.atom_counts = atom_counts.data();
    b.charge = charge.data();
    b.lj_type = lj_type.data();
    b.lj_type_counts = lj_types.data();
    b.pb_radii = radii.data();
    b.gb_screen = screen.data();
    b.bond_counts = bond_counts.data();
    b.bond_atoms = bonds.data();
    b.bond_k = bond_k.data();
    b.bond_r0 = bond_r0.data();
    b.angle_counts = angle_counts.data();
    b.angle_atoms = angles.data();
    b.angle_k = angle_k.data();
    b.angle_theta = angle_t.data();
    b.dihedral_counts = dihe_counts.data();
    b.dihedral_atoms = dihedrals.data();
    b.dihe_amp = amp.data();
    b.dihe_phase = phase.data();
    b.dihe_freq = freq.data();
    b.dihe_improper = improper.data();
    b.dihe_scee = scee.data();
    b.dihe_scnb = scnb.data();
    b.igb = STORMM_IGB_OBC2;
    b.dielectric = 78.5;
    b.coulomb_constant = 332.0522173;
    b.lj_combining = combining.data();
    b.lj_sigma = sigma.data();
    b.lj_epsilon = epsilon.data();
    return b;
  }
};

// Device arrays of one caller: 12 word arrays and 12 overflow arrays of padded length, a
// ScoreCard row block and a line-minimization record.
struct CallerArrays {
  std::vector<void *> owned;
  stormm_ps_writer psw;
  stormm_scorecard_writer scw;
  stormm_linmin_writer lmw;

  template <typename T> T *alloc(size_t n) {
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess || hipMemset(p, 0, n * sizeof(T)) != hipSuccess) {
      return nullptr;
    }
    owned.push_back(p);
    return static_cast<T *>(p);
  }
  ~CallerArrays() {
    for (void *p : owned) {
      (void)hipFree(p);
    }
  }
};

int setup(CallerArrays &c, const stormm_synthesis_info &info, const int32_t *starts_h,
          const int32_t *counts_h, int gpos, int frc, int nrg, double dx0) {
  const size_t np = static_cast<size_t>(info.padded_atom_count);
  const size_t ns = static_cast<size_t>(info.system_count);
  std::memset(&c.psw, 0, sizeof(c.psw));
  c.psw.system_count = info.system_count;
  c.psw.gpos_bits = gpos;
  c.psw.frc_bits = frc;
  c.psw.gpos_scale = std::ldexp(1.0, gpos);
  c.psw.inv_gpos_scale = std::ldexp(1.0, -gpos);
  c.psw.frc_scale = std::ldexp(1.0, frc);
  c.psw.inv_frc_scale = std::ldexp(1.0, -frc);
  int32_t *starts = c.alloc<int32_t>(ns), *counts = c.alloc<int32_t>(ns);
  HIPCHECK(hipMemcpy(starts, starts_h, ns * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(counts, counts_h, ns * sizeof(int32_t), hipMemcpyHostToDevice));
  c.psw.atom_starts = starts;
  c.psw.atom_counts = counts;
  int64_t **words[12] = {&c.psw.xcrd, &c.psw.ycrd, &c.psw.zcrd, &c.psw.xfrc, &c.psw.yfrc,
                         &c.psw.zfrc, &c.psw.xprv, &c.psw.yprv, &c.psw.zprv, &c.psw.x_cg_temp,
                         &c.psw.y_cg_temp, &c.psw.z_cg_temp};
  int32_t **ovrf[12] = {&c.psw.xcrd_ovrf, &c.psw.ycrd_ovrf, &c.psw.zcrd_ovrf,
                        &c.psw.xfrc_ovrf, &c.psw.yfrc_ovrf, &c.psw.zfrc_ovrf,
                        &c.psw.xprv_ovrf, &c.psw.yprv_ovrf, &c.psw.zprv_ovrf,
                        &c.psw.x_cg_temp_ovrf, &c.psw.y_cg_temp_ovrf, &c.psw.z_cg_temp_ovrf};
  for (int i = 0; i < 12; i++) {
    *words[i] = c.alloc<int64_t>(np);
    *ovrf[i] = c.alloc<int32_t>(np);
  }
  std::memset(&c.scw, 0, sizeof(c.scw));
  c.scw.system_count = info.system_count;
  c.scw.data_stride = STORMM_SCORECARD_STRIDE;
  c.scw.nrg_bits = nrg;
  c.scw.instantaneous_accumulators = c.alloc<int64_t>(ns * STORMM_SCORECARD_STRIDE);
  double *line = c.alloc<double>(9 * ns);
  std::vector<double> l0(ns, dx0);
  HIPCHECK(hipMemcpy(line, l0.data(), ns * sizeof(double), hipMemcpyHostToDevice));
  c.lmw.system_count = info.system_count;
  double **f[9] = {&c.lmw.l_move, &c.lmw.s_move, &c.lmw.mfac_a, &c.lmw.mfac_b, &c.lmw.mfac_c,
                   &c.lmw.nrg_a,  &c.lmw.nrg_b,  &c.lmw.nrg_c,  &c.lmw.nrg_d};
  for (int j = 0; j < 9; j++) {
    *f[j] = line + j * ns;
  }
  return 0;
}

// Scatters per-atom values ([atoms][3], system after system) to the padded device arrays.
template <typename T>
int scatter(const std::vector<T> &v, const int32_t *starts, const int32_t *counts, int nsys,
            size_t np, T *dx, T *dy, T *dz) {
  std::vector<T> a(np, 0), b(np, 0), c(np, 0);
  size_t k = 0;
  for (int s = 0; s < nsys; s++) {
    for (int i = 0; i < counts[s]; i++, k++) {
      a[starts[s] + i] = v[3 * k];
      b[starts[s] + i] = v[3 * k + 1];
      c[starts[s] + i] = v[3 * k + 2];
    }
  }
  HIPCHECK(hipMemcpy(dx, a.data(), np * sizeof(T), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(dy, b.data(), np * sizeof(T), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(dz, c.data(), np * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

template <typename T>
int gather(std::vector<T> &v, const int32_t *starts, const int32_t *counts, int nsys, size_t np,
           const T *dx, const T *dy, const T *dz) {
  std::vector<T> a(np), b(np), c(np);
  HIPCHECK(hipDeviceSynchronize());
  HIPCHECK(hipMemcpy(a.data(), dx, np * sizeof(T), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(b.data(), dy, np * sizeof(T), hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(c.data(), dz, np * sizeof(T), hipMemcpyDeviceToHost));
  v.clear();
  for (int s = 0; s < nsys; s++) {
    for (int i = 0; i < counts[s]; i++) {
      v.push_back(a[starts[s] + i]);
      v.push_back(b[starts[s] + i]);
      v.push_back(c[starts[s] + i]);
    }
  }
  return 0;
}

// One case: evaluation and a 12-cycle minimization (per-step launches and graph replay) on the
// handles and on the caller's arrays; every word must agree.
int run_case(const stormm_topology_batch &batch, const std::vector<double> &xyz, int prec,
             int gpos, int frc, int nrg) {
  stormm_synthesis *ag = nullptr;
  CHECK(stormm_synthesis_create(&batch, nullptr, &ag));
  stormm_synthesis_info info;
  CHECK(stormm_synthesis_get_info(ag, &info));
  const int ns = info.system_count;
  const size_t np = static_cast<size_t>(info.padded_atom_count);
  std::vector<int32_t> starts(ns), counts(ns);
  CHECK(stormm_synthesis_atom_layout(ag, starts.data(), counts.data()));
  size_t natom = 0;
  for (int s = 0; s < ns; s++) {
    natom += counts[s];
  }
  int mismatches = 0;
  for (int mode = -1; mode <= 2; mode += (mode < 0 ? 1 : 2)) {
    // Handles.
    stormm_phasespace *ps = nullptr;
    stormm_scorecard *sc = nullptr;
    stormm_line_min *lm = nullptr;
    CHECK(stormm_phasespace_create(ag, gpos, frc, &ps));
    CHECK(stormm_scorecard_create(ns, nrg, &sc));
    CHECK(stormm_line_min_create(ns, 0.01, &lm));
    CHECK(stormm_phasespace_upload_coordinates(ps, xyz.data()));
    std::vector<int64_t> c0(3 * natom), f0(3 * natom);
    std::vector<int32_t> c0o(3 * natom), f0o(3 * natom);
    CHECK(stormm_phasespace_download_fixed_split(ps, c0.data(), c0o.data(), nullptr, nullptr));
    stormm_minimize_controls mc = {12, 5, 0.01, 0.0, 3, 0.2, 0.5};
    if (mode < 0) {
      CHECK(stormm_phasespace_initialize_forces(ps, nullptr));
      CHECK(stormm_scorecard_initialize(sc, nullptr));
      CHECK(stormm_launch_mm_evaluation(prec, ag, nullptr, ps, sc, 1, 1, 0.0, 0.0, nullptr));
    }
    else {
      CHECK(stormm_launch_minimization(prec, ag, ps, &mc, sc, lm, mode, nullptr));
    }
    std::vector<int64_t> c1(3 * natom), f1(3 * natom), e1(ns * STORMM_SCORECARD_STRIDE);
    std::vector<int32_t> c1o(3 * natom), f1o(3 * natom);
    CHECK(stormm_phasespace_download_fixed_split(ps, c1.data(), c1o.data(), f1.data(),
                                                 f1o.data()));
    CHECK(stormm_scorecard_download(sc, e1.data()));
    // The caller's arrays, starting from the same fixed-precision coordinates.
    CallerArrays mine;
    if (setup(mine, info, starts.data(), counts.data(), gpos, frc, nrg, 0.01) != 0 ||
        scatter(c0, starts.data(), counts.data(), ns, np, mine.psw.xcrd, mine.psw.ycrd,
                mine.psw.zcrd) != 0 ||
        scatter(c0o, starts.data(), counts.data(), ns, np, mine.psw.xcrd_ovrf,
                mine.psw.ycrd_ovrf, mine.psw.zcrd_ovrf) != 0) {
      return 1;
    }
    CHECK(stormm_check_ps_writer(ag, &mine.psw));
    if (mode < 0) {
      CHECK(stormm_ps_writer_initialize_forces(&mine.psw, info.padded_atom_count, nullptr));
      CHECK(stormm_scorecard_writer_initialize(&mine.scw, nullptr));
      CHECK(stormm_launch_mm_evaluation_kit(prec, ag, nullptr, &mine.psw, &mine.scw, 1, 1, 0.0,
                                            0.0, nullptr));
    }
    else {
      CHECK(stormm_launch_minimization_kit(prec, ag, &mine.psw, &mc, &mine.scw, &mine.lmw, mode,
                                           nullptr));
    }
    std::vector<int64_t> c2, f2, e2(ns * STORMM_SCORECARD_STRIDE);
    std::vector<int32_t> c2o, f2o;
    if (gather(c2, starts.data(), counts.data(), ns, np, mine.psw.xcrd, mine.psw.ycrd,
               mine.psw.zcrd) != 0 ||
        gather(c2o, starts.data(), counts.data(), ns, np, mine.psw.xcrd_ovrf,
               mine.psw.ycrd_ovrf, mine.psw.zcrd_ovrf) != 0 ||
        gather(f2, starts.data(), counts.data(), ns, np, mine.psw.xfrc, mine.psw.yfrc,
               mine.psw.zfrc) != 0 ||
        gather(f2o, starts.data(), counts.data(), ns, np, mine.psw.xfrc_ovrf,
               mine.psw.yfrc_ovrf, mine.psw.zfrc_ovrf) != 0) {
      return 1;
    }
    HIPCHECK(hipMemcpy(e2.data(), mine.scw.instantaneous_accumulators,
                       e2.size() * sizeof(int64_t), hipMemcpyDeviceToHost));
    // Compare values (word and overflow word; the handle reports canonical splits, so compare
    // value = ovrf 2^63 + word).
    auto value = [](int64_t w, int32_t o) {
      return (static_cast<__int128>(o) << 63) + static_cast<__int128>(w);
    };
    const bool wide_f = frc > 40;
    const bool wide_c = gpos > 38;
    int bad = 0;
    for (size_t k = 0; k < 3 * natom; k++) {
      bad += value(c1[k], c1o[k]) != value(c2[k], wide_c ? c2o[k] : 0);
      bad += value(f1[k], f1o[k]) != value(f2[k], wide_f ? f2o[k] : 0);
    }
    for (size_t k = 0; k < e1.size(); k++) {
      bad += e1[k] != e2[k];
    }
    std::printf("%s %d/%d bits, %s: %d systems, %zu atoms, %d mismatches\n",
                prec == STORMM_PREC_DOUBLE ? "double" : "single", gpos, frc,
                mode < 0 ? "evaluation" : (mode == 0 ? "minimization (per-step launches)"
                                                     : "minimization (graph replay)"),
                ns, natom, bad);
    mismatches += bad;
    stormm_line_min_destroy(lm);
    stormm_scorecard_destroy(sc);
    stormm_phasespace_destroy(ps);
  }
  stormm_synthesis_destroy(ag);
  return mismatches == 0 ? 0 : 1;
}

}  // namespace

int main() {
  std::mt19937 rng(91);
  Molecules m;
  for (int n : {12, 30, 47, 64, 90}) {
    m.add(n, rng);
  }
  const stormm_topology_batch batch = m.batch();
  int rc = 0;
  rc |= run_case(batch, m.xyz, STORMM_PREC_SINGLE, 32, 23, 26);
  rc |= run_case(batch, m.xyz, STORMM_PREC_DOUBLE, 40, 40, 32);
  rc |= run_case(batch, m.xyz, STORMM_PREC_DOUBLE, 60, 60, 32);
  std::printf("%s\n", rc == 0 ? "kit boundary ok" : "kit boundary FAILED");
  return rc;
}
