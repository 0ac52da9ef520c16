/*
 * stormm_oracle.h -- CPU double-precision restatement of the reference's molecular-mechanics
 * energy/force path (valence terms, non-periodic Coulomb + Lennard-Jones, HCT / OBC-I / OBC-II
 * Generalized Born) and of the GPU minimizer's conjugate-gradient and line-search semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only the test suite, the smoke entry
 * point and the CPU-baseline leg of the benchmark load it.  The product path (stormm_amd) never
 * links, imports or calls it.
 *
 * Every routine names the reference file:line it restates.  Paths are relative to the reference
 * repository root (suobset/stormm).
 */
#ifndef STORMM_ORACLE_H
#define STORMM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Energy slots follow the reference StateVariable enumeration
 * (src/Potential/energy_enumerators.h:36-75). */
enum {
  ORC_BOND = 0, ORC_ANGLE = 1, ORC_PROPER_DIHEDRAL = 2, ORC_IMPROPER_DIHEDRAL = 3,
  ORC_UREY_BRADLEY = 4, ORC_CHARMM_IMPROPER = 5, ORC_CMAP = 6, ORC_VDW = 7, ORC_VDW_ONE_FOUR = 8,
  ORC_ELECTROSTATIC = 9, ORC_ELEC_ONE_FOUR = 10, ORC_GENERALIZED_BORN = 11, ORC_RESTRAINT = 12,
  ORC_SLOTS = 32
};

/* Term masks selecting which terms contribute forces in orc_eval_terms. */
enum {
  ORC_TERM_BOND = 1u, ORC_TERM_ANGLE = 2u, ORC_TERM_DIHEDRAL = 4u, ORC_TERM_QQ14 = 8u,
  ORC_TERM_LJ14 = 16u, ORC_TERM_QQ = 32u, ORC_TERM_LJ = 64u, ORC_TERM_GB = 128u,
  ORC_TERM_ALL = 255u
};

/* Implicit solvent models (AMBER igb numbering). */
enum { ORC_IGB_NONE = 0, ORC_IGB_HCT = 1, ORC_IGB_OBC = 2, ORC_IGB_OBC2 = 5 };

/* One flat-bottom restraint, the fields of RestraintApparatus / RestraintKit
 * (src/Restraints/restraint_apparatus.h): kind 1 positional (atoms[0] to a target point),
 * 2 distance (atoms 0-1), 3 angle (0-1-2), 4 dihedral (0-1-2-3).  k[0..1] are k2 and k3, r[0..3]
 * are r1..r4 (A or radians); values slide from init_* to final_* between the two steps. */
typedef struct {
  int kind;
  int atoms[4];
  int init_step, final_step;
  double init_k[2], final_k[2];
  double init_r[4], final_r[4];
  double init_target[3], final_target[3];
} orc_restraint;

enum { ORC_RST_POSITION = 1, ORC_RST_DISTANCE = 2, ORC_RST_ANGLE = 3, ORC_RST_DIHEDRAL = 4 };

typedef struct {
  int natom;
  double coulomb;                 /* Coulomb constant (kcal/mol A / e^2)                        */
  const double *charge;           /* [natom] partial charges (e)                                */
  int n_lj_types;
  const int *lj_idx;              /* [natom] Lennard-Jones type index                          */
  const double *lja, *ljb;        /* [ntypes * ntypes] A and B coefficients                     */
  const double *lja14, *ljb14;    /* [ntypes * ntypes] 1:4 A and B coefficients                 */
  const unsigned char *excl;      /* [natom * natom] 1 = pair excluded from Coulomb / LJ        */
  int nbond;
  const int *bond_ij;             /* [2 * nbond]                                                */
  const double *bond_keq, *bond_leq;
  int nangl;
  const int *angl_ijk;            /* [3 * nangl]                                                */
  const double *angl_keq, *angl_theta;
  int ndihe;
  const int *dihe_ijkl;           /* [4 * ndihe]                                                */
  const double *dihe_amp, *dihe_phase, *dihe_freq;
  const int *dihe_improper;       /* [ndihe] 0 = proper, 1 = improper                           */
  const double *dihe_scee;        /* [ndihe] electrostatic 1:4 divisor, <= 0 means no 1:4 pair  */
  const double *dihe_scnb;        /* [ndihe] Lennard-Jones 1:4 divisor                          */
  int ninfr14;
  const int *infr14_il;           /* [2 * ninfr14] inferred 1:4 pairs                           */
  const double *infr14_scee, *infr14_scnb;
  int igb;
  double dielectric, kappa, gb_offset;
  const double *pb_radii, *gb_screen, *gb_alpha, *gb_beta, *gb_gamma;
  int nrst;                       /* restraints (may be 0)                                      */
  const orc_restraint *rst;
} orc_system;

/* Energies of one evaluation: plain double sums and the fixed-precision accumulators that the
 * reference ScoreCard holds (llround(term * 2^nrg_bits) per thread-equivalent block). */
typedef struct {
  double nrg[ORC_SLOTS];
  long long acc[ORC_SLOTS];
} orc_energy;

/* Non-bonded Coulomb + Lennard-Jones, src/Potential/nonbonded_potential.tpp:9-425. */
void orc_nonbonded(const orc_system *s, const double *x, const double *y, const double *z,
                   double *fx, double *fy, double *fz, int eval_force, double nrg_scale,
                   orc_energy *e);

/* Generalized Born, src/Potential/nonbonded_potential.tpp:449-1292.  psi, radii and sumdeijda
 * are per-atom work arrays (may be NULL, in which case they are allocated internally). */
void orc_gb(const orc_system *s, const double *x, const double *y, const double *z, double *fx,
            double *fy, double *fz, int eval_force, double nrg_scale, orc_energy *e, double *psi,
            double *radii, double *sumdeijda);

/* Valence terms: bonds, angles, cosine dihedrals, attenuated 1:4 pairs,
 * src/Potential/valence_potential.tpp:9-96,170-287,359-578,1288-1523. */
void orc_valence(const orc_system *s, const double *x, const double *y, const double *z,
                 double *fx, double *fy, double *fz, int eval_force, double nrg_scale,
                 orc_energy *e);

/* Composite evaluation in the order of evalRestrainedMMGB,
 * src/MolecularMechanics/mm_evaluation.tpp:630-652 (non-bonded, valence, then GB).  Forces are
 * added to fx, fy, fz when eval_force != 0 (the caller zeroes them). */
void orc_eval_mm(const orc_system *s, const double *x, const double *y, const double *z,
                 double *fx, double *fy, double *fz, int eval_force, int nrg_bits, orc_energy *e);

/* orc_eval_mm with clash damping of the Coulomb, Lennard-Jones and 1:4 terms (both parameters
 * zero gives the plain potentials): src/Potential/soft_core_potentials.tpp:32-55,145-188 as
 * called from nonbonded_potential.tpp:108-172,244-311 and valence_potential.tpp:1337-1360. */
void orc_eval_mm_clash(const orc_system *s, const double *x, const double *y, const double *z,
                       double *fx, double *fy, double *fz, int eval_force, int nrg_bits,
                       double clash_distance, double clash_ratio, orc_energy *e);

/* orc_eval_mm_clash plus the system's restraints evaluated at step_number
 * (src/Potential/valence_potential.tpp:2090-2160, src/Restraints/restraint_util.tpp). */
void orc_eval_mm_ext(const orc_system *s, const double *x, const double *y, const double *z,
                     double *fx, double *fy, double *fz, int eval_force, int nrg_bits,
                     double clash_distance, double clash_ratio, int step_number, orc_energy *e);

/* Evaluate all energies, accumulating forces only from the terms named in force_mask (the
 * per-term force flags of the reference's evaluators, e.g. nonbonded_potential.tpp:13-14). */
void orc_eval_terms(const orc_system *s, const double *x, const double *y, const double *z,
                    double *fx, double *fy, double *fz, unsigned force_mask, int nrg_bits,
                    orc_energy *e);

/* Batched evaluation of many systems on nthreads POSIX threads (disjoint system ranges), used as
 * the timed CPU baseline.  coordinates are concatenated per system (offsets[i] .. offsets[i] +
 * natom).  Returns wall seconds for `repeats` passes over the batch. */
double orc_eval_batch(int nsys, const orc_system *systems, const long long *offsets,
                      const double *x, const double *y, const double *z, double *fx, double *fy,
                      double *fz, int eval_force, int nrg_bits, int repeats, int nthreads,
                      double *total_energy);

/* Cubic line-search move selection of the GPU line minimizer (move 3),
 * src/MolecularMechanics/line_movement.cui:129-338.  Returns the move factor and writes the
 * multiplier applied to the baseline move length. */
double orc_cubic_move(const double evec[4], const double mvec[4], double *new_baseline_move,
                      double coefs_out[4]);

/* Minimization with the GPU semantics of launchMinimization
 * (src/MolecularMechanics/hpc_minimization.cu:262-405): conjugate gradient per
 * src/Math/hpc_reduction.cu:39-244 and src/Math/conjugate_gradient.cui:259-370, line moves per
 * src/MolecularMechanics/line_movement.cui.  Coordinates are fixed-precision integers at
 * gpos_bits; forces are fixed-precision at frc_bits.  dp_mode selects the double-precision
 * model's rounding of the normalized move (doubleToInt95) instead of the single-precision
 * model's truncation.  The first clash_cycles cycles evaluate with clash damping whose floor
 * (clash_r0) and ratio shrink linearly to zero (hpc_minimization.cu:269-274).  Restraints see
 * the reference's step counters: the force evaluation of cycle i step i, its three energy
 * evaluations steps 3i, 3i + 1, 3i + 2, the final evaluation step 3 nsteps.  energy_trace
 * receives the energy (kcal/mol) at the start of each step and, in its last slot, the final
 * energy (length nsteps + 1); may be NULL. */
void orc_minimize(const orc_system *s, long long *xcrd, long long *ycrd, long long *zcrd,
                  int gpos_bits, int frc_bits, int nrg_bits, int nsteps, int sd_cycles,
                  double dx0, int clash_cycles, double clash_r0, double clash_ratio,
                  int dp_mode, double *energy_trace);

/* orc_minimize, also writing for each cycle ORC_LINE_RECORD_LENGTH doubles of the line search
 * (the LineMinimization record, src/MolecularMechanics/line_minimization.h:85-221): the four
 * energies nrg_a..nrg_d, the move factors of moves 1 and 2 (mfac_b, mfac_c), the move factor of
 * the cubic-fit move, the baseline move length for the next cycle and the conjugate-gradient
 * gamma used this cycle.  line_record may be NULL. */
#define ORC_LINE_RECORD_LENGTH 9
void orc_minimize_record(const orc_system *s, long long *xcrd, long long *ycrd, long long *zcrd,
                         int gpos_bits, int frc_bits, int nrg_bits, int nsteps, int sd_cycles,
                         double dx0, int clash_cycles, double clash_r0, double clash_ratio,
                         int dp_mode, double *energy_trace, double *line_record);

/* Test hook: make the minimizer commit one deliberate mistake (or none, code 0) so that a test
 * can show that its checks fail on that mistake.  Not thread-safe; tests only. */
#define ORC_MISTAKE_NONE 0
#define ORC_MISTAKE_FLETCHER_REEVES 1      /* gamma = |f|^2 / |f_prev|^2                     */
#define ORC_MISTAKE_NO_CONJUGATION 2       /* gamma = 0 every cycle                          */
#define ORC_MISTAKE_GAMMA_SIGN 3           /* gamma with its sign flipped                    */
#define ORC_MISTAKE_FIRST_GROWTH 4         /* 1.05 in place of 1.04 after an improving move 1 */
#define ORC_MISTAKE_FIRST_SHRINK 5         /* 0.75 in place of 0.8 after a failing move 1     */
#define ORC_MISTAKE_SECOND_GROWTH 6        /* 1.05 in place of 1.03 after an improving move 2 */
#define ORC_MISTAKE_NORMALIZATION_ROUNDING 7 /* round-to-nearest and truncation swapped       */
void orc_plant_mistake(int code);

#ifdef __cplusplus
}
#endif
#endif
