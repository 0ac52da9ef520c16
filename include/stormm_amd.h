/*
 * stormm_amd.h -- C-ABI of the MI355X batched molecular-mechanics energy/force path.
 *
 * The entry points replace the reference's launch surface for this path (paths relative to the
 * reference repository root, suobset/stormm):
 *
 *   stormm_launch_nonbonded          launchNonbonded        src/Potential/hpc_nonbonded_potential.h:142-157
 *   stormm_launch_valence            launchValence          src/Potential/hpc_valence_potential.h:356-378
 *   stormm_launch_mm_evaluation      launchNonbonded + launchValence back to back, as driven by
 *                                    test/Synthesis/test_hpc_synthesis.cu:106,143
 *   stormm_launch_conjugate_gradient launchConjugateGradient src/Math/hpc_reduction.h:55-63
 *   stormm_launch_line_advance       launchLineAdvance      src/MolecularMechanics/hpc_minimization.h:63-70
 *   stormm_launch_minimization       launchMinimization     src/MolecularMechanics/hpc_minimization.h:122-151
 *   stormm_phasespace_initialize_forces   PhaseSpaceSynthesis::initializeForces
 *                                         src/Synthesis/phasespace_synthesis.h:787-797
 *   stormm_phasespace_prime_conjugate_gradient  PhaseSpaceSynthesis::primeConjugateGradientCalculation
 *                                         src/Synthesis/phasespace_synthesis.h:799-808
 *   stormm_scorecard_initialize      ScoreCard::initialize  src/Potential/scorecard.h:212-222
 *   stormm_scorecard_commit          ScoreCard::commit      src/Potential/scorecard.h:248-262
 *
 * Objects (AtomGraphSynthesis, PhaseSpaceSynthesis, ScoreCard, MolecularMechanicsControls,
 * LineMinimization) are opaque handles owning device memory.  Their device layouts follow the
 * reference abstracts: fixed-precision int64 coordinates and forces per atom with each system's
 * atoms starting on a 32-atom boundary (PsSynthesisWriter, src/Synthesis/phasespace_synthesis.h:
 * 87-195), ScoreCard int64 accumulators with a stride of 32 slots per system in StateVariable
 * order (src/Potential/scorecard.h:64-95, src/Potential/energy_enumerators.h:36-75).
 *
 * Every function returns 0 on success and a negative STORMM_ERR_* code on failure;
 * stormm_last_error() returns a description of the last failure on the calling thread.
 * Streams are hipStream_t passed as void* (NULL = the default stream).
 */
#ifndef STORMM_AMD_H
#define STORMM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STORMM_OK 0
#define STORMM_ERR_INVALID_ARGUMENT (-1)
#define STORMM_ERR_HIP (-2)
#define STORMM_ERR_CAPACITY (-3)
#define STORMM_ERR_STATE (-4)

/* Precision models, src/Constants/behavior.h (PrecisionModel). */
#define STORMM_PREC_DOUBLE 0
#define STORMM_PREC_SINGLE 1

/* Implicit solvent models (AMBER igb numbers). */
#define STORMM_IGB_NONE 0
#define STORMM_IGB_HCT 1
#define STORMM_IGB_OBC 2
#define STORMM_IGB_OBC2 5

/* ScoreCard slots in StateVariable order, src/Potential/energy_enumerators.h:36-75. */
#define STORMM_NRG_BOND 0
#define STORMM_NRG_ANGLE 1
#define STORMM_NRG_PROPER_DIHEDRAL 2
#define STORMM_NRG_IMPROPER_DIHEDRAL 3
#define STORMM_NRG_VDW 7
#define STORMM_NRG_VDW_ONE_FOUR 8
#define STORMM_NRG_ELECTROSTATIC 9
#define STORMM_NRG_ELEC_ONE_FOUR 10
#define STORMM_NRG_GENERALIZED_BORN 11
#define STORMM_NRG_RESTRAINT 12
#define STORMM_NRG_POTENTIAL_ENERGY 28
#define STORMM_SCORECARD_STRIDE 32

/* Flat-bottom restraints: one record holds the fields of RestraintApparatus / RestraintKit
 * (src/Restraints/restraint_apparatus.h) for one restraint.  Positional restraints hold atom 0
 * to a target point, distance restraints atoms 0-1, angle restraints 0-1-2 and dihedral
 * restraints 0-1-2-3 (atom indices local to the system).  k[0], k[1] are the stiffnesses k2, k3
 * and r[0..3] the bounds r1 <= r2 <= r3 <= r4 (A, or radians for angles and dihedrals; dihedral
 * bounds span at most 2 pi): the penalty is quadratic between r1 and r2 and between r3 and r4,
 * zero between r2 and r3, linear outside.  Parameters slide linearly from init_* to final_*
 * between init_step and final_step of the controls' step counter; before init_step the
 * restraint is off (src/Restraints/restraint_util.tpp). */
#define STORMM_RESTRAINT_POSITION 1
#define STORMM_RESTRAINT_DISTANCE 2
#define STORMM_RESTRAINT_ANGLE 3
#define STORMM_RESTRAINT_DIHEDRAL 4
typedef struct {
  int32_t kind;
  int32_t atoms[4];
  int32_t init_step, final_step;
  double init_k[2], final_k[2];
  double init_r[4], final_r[4];
  double init_target[3], final_target[3];   /* positional restraints (A)                       */
} stormm_restraint;

/* Many molecules' parameters, concatenated system after system.  Atom indices inside the term
 * arrays are local to their system.  This is the input the reference assembles from a list of
 * AtomGraph objects in the AtomGraphSynthesis constructor
 * (src/Synthesis/ag_synthesis_constructors.cpp:11-468). */
typedef struct {
  int32_t system_count;
  const int32_t *atom_counts;      /* [nsys]                                                    */
  const double *charge;            /* [sum atoms] e                                             */
  const int32_t *lj_type;          /* [sum atoms] type index local to the system's table        */
  const int32_t *lj_type_counts;   /* [nsys] T                                                  */
  const double *lja, *ljb;         /* [sum T^2 over table systems] A / B, row-major (ti*T + tj)  */
  const double *lja14, *ljb14;     /* [sum T^2 over table systems] 1:4 A / B                     */
  const double *pb_radii;          /* [sum atoms] intrinsic radii (A)                           */
  const double *gb_screen;         /* [sum atoms] screening factors                             */
  const int32_t *bond_counts;      /* [nsys]                                                    */
  const int32_t *bond_atoms;       /* [2 * sum bonds]                                           */
  const double *bond_k, *bond_r0;
  const int32_t *angle_counts;
  const int32_t *angle_atoms;      /* [3 * sum angles]                                          */
  const double *angle_k, *angle_theta;
  const int32_t *dihedral_counts;
  const int32_t *dihedral_atoms;   /* [4 * sum dihedrals]                                       */
  const double *dihe_amp, *dihe_phase, *dihe_freq;
  const int32_t *dihe_improper;    /* 0 proper, 1 improper                                      */
  const double *dihe_scee;         /* 1:4 electrostatic divisor; <= 0: no 1:4 pair on this term */
  const double *dihe_scnb;         /* 1:4 Lennard-Jones divisor                                 */
  const int32_t *exclusion_counts; /* [nsys] number of excluded (i < j) pairs; null: the pairs   */
  const int32_t *exclusion_pairs;  /* within three bonds [2 * sum exclusions]                   */
  int32_t igb;                     /* one implicit-solvent model for the whole synthesis        */
  double dielectric;
  double salt_concentration;
  double coulomb_constant;
  /* Optional compact Lennard-Jones form (no counterpart in the reference, whose AtomGraph always
   * holds the T x T tables).  lj_combining[s] = 1: system s follows Lorentz-Berthelot combining
   * of per-type sigma / epsilon given in lj_sigma / lj_epsilon ([sum T over such systems]), the
   * 1:4 pairs use the same coefficients, and the system has no block in lja..ljb14.  Null
   * lj_combining: every system sends its tables. */
  const int32_t *lj_combining;     /* [nsys] or null                                            */
  const double *lj_sigma;          /* A                                                         */
  const double *lj_epsilon;        /* kcal/mol                                                  */
  /* Restraints of each system, the reference's per-system RestraintApparatus list handed to the
   * synthesis; they feed the RESTRAINT ScoreCard slot.  Null restraint_counts: none. */
  const int32_t *restraint_counts; /* [nsys] or null                                            */
  const stormm_restraint *restraints; /* [sum counts], system after system                      */
} stormm_topology_batch;

/* Work-unit shaping, chosen by the autotuner (or fixed by the caller).  Zero = default. */
typedef struct {
  int32_t wu_atom_capacity;        /* atoms staged in LDS per work unit (<= 512)                */
  int32_t target_wu_per_cu;        /* work units per compute unit to aim for when packing       */
  int32_t block_threads;           /* threads per work-unit block: 256, 512 (default) or 1024   */
} stormm_launch_shape;

typedef struct stormm_synthesis stormm_synthesis;
typedef struct stormm_phasespace stormm_phasespace;
typedef struct stormm_scorecard stormm_scorecard;
typedef struct stormm_mm_controls stormm_mm_controls;
typedef struct stormm_line_min stormm_line_min;

typedef struct {
  int32_t system_count;
  int32_t padded_atom_count;       /* length of the per-atom device arrays                      */
  int32_t work_unit_count;
  int32_t task_count;
  int32_t max_wu_atoms;
  int32_t lj_table_systems;        /* systems using the stored LJ table (not a combining rule)  */
  int64_t pair_count;              /* sum over systems of N (N - 1) / 2                         */
  int64_t valence_term_count;
} stormm_synthesis_info;

/* Device pointers of a PhaseSpaceSynthesis, the PsSynthesisWriter layout
 * (src/Synthesis/phasespace_synthesis.h:87-195): per-atom fixed-precision arrays, each system
 * starting at atom_starts[s].  Values above the reference's plain 64-bit limits are split into a
 * 64-bit word and a 32-bit overflow word, value = ovrf * 2^63 + word
 * (src/Numerics/accumulation.cui:121-403): coordinates when gpos_bits > 38, forces, previous
 * forces and conjugate-gradient directions when frc_bits > 40.  Overflow arrays are read and
 * written only then (they may be NULL otherwise).  The reference keeps the previous forces in the
 * alternate-coordinate arrays (xalt) and the conjugate-gradient directions in the velocity
 * arrays (xvel) of the writer (src/Math/reduction_abstracts.cpp:25-39). */
typedef struct {
  int32_t system_count;
  int32_t gpos_bits, frc_bits;
  double gpos_scale, inv_gpos_scale, frc_scale, inv_frc_scale;
  const int32_t *atom_starts, *atom_counts;
  int64_t *xcrd, *ycrd, *zcrd;
  int64_t *xfrc, *yfrc, *zfrc;
  int64_t *xprv, *yprv, *zprv;          /* prior-step forces for conjugate gradient            */
  int64_t *x_cg_temp, *y_cg_temp, *z_cg_temp;
  int32_t *xcrd_ovrf, *ycrd_ovrf, *zcrd_ovrf;
  int32_t *xfrc_ovrf, *yfrc_ovrf, *zfrc_ovrf;
  int32_t *xprv_ovrf, *yprv_ovrf, *zprv_ovrf;
  int32_t *x_cg_temp_ovrf, *y_cg_temp_ovrf, *z_cg_temp_ovrf;
} stormm_ps_writer;

/* ScoreCardWriter layout (src/Potential/scorecard.h:64-95): instantaneous accumulators, one row
 * of data_stride (32) int64 slots per system in StateVariable order, at nrg_bits, and the
 * energy history: running sums and sums of squares of the committed samples (kcal/mol) and the
 * time series, sample after sample, each sample one row per system.  A caller without history
 * leaves the last five fields zero. */
typedef struct {
  int32_t system_count;
  int32_t data_stride;
  int32_t nrg_bits;
  int64_t *instantaneous_accumulators;
  int32_t sampled_step_count;      /* samples already in time_series                          */
  int32_t sample_capacity;         /* samples time_series holds                               */
  double *running_accumulators;    /* [system_count][32], or NULL                             */
  double *squared_accumulators;    /* [system_count][32], NULL exactly when running is NULL   */
  int64_t *time_series;            /* [sample_capacity][system_count][32], or NULL            */
} stormm_scorecard_writer;

/* LinMinWriter layout (src/MolecularMechanics/line_minimization.h:15-50): per-system line-search
 * state in double precision. */
typedef struct {
  int32_t system_count;
  double *l_move, *s_move, *mfac_a, *mfac_b, *mfac_c, *nrg_a, *nrg_b, *nrg_c, *nrg_d;
} stormm_linmin_writer;

/* The fields of MMControlKit (src/MolecularMechanics/mm_controls.h:52-93) this path reads: the
 * step counter (restraint mixing, steepest descent vs conjugate gradient) and the number of
 * steepest-descent cycles.  The work-unit progress counters are not needed (each block owns one
 * work unit). */
typedef struct {
  int32_t step;
  int32_t sd_cycles;
} stormm_mm_kit;

/* Minimization controls, the subset of MinimizeControls this path reads
 * (src/Namelists/nml_minimize.h:21-30).  The reference's defaults are maxcyc 200, ncyc 50,
 * cdcyc 25, dx0 0.01, clash_r0 0.2, clash_vdw_ratio 0.5.  During the first cdcyc cycles the
 * evaluations damp clashes with a floor and ratio shrinking linearly to zero
 * (src/MolecularMechanics/hpc_minimization.cu:269-281); 0 <= cdcyc <= maxcyc. */
typedef struct {
  int32_t total_cycles;            /* maxcyc                                                    */
  int32_t steepest_descent_cycles; /* ncyc                                                      */
  double initial_step;             /* dx0 (A)                                                   */
  double convergence_rms_force;    /* stop a system once its RMS force <= this (0 = never)      */
  int32_t clash_damping_cycles;    /* cdcyc                                                     */
  double clash_minimum_distance;   /* clash_r0 (A): Coulomb is softened inside this distance     */
  double clash_vdw_ratio;          /* clash_vdw_ratio: LJ is softened inside ratio * sigma_ij    */
  int32_t diagnostic_print_frequency; /* ntpr: commit the ScoreCard every ntpr cycles (0: never) */
} stormm_minimize_controls;

const char *stormm_last_error(void);
int stormm_device_info(int32_t *cu_count, int32_t *lds_bytes_per_cu, char *name, int32_t name_len);
/* Wait for all work on the device (cudaDeviceSynchronize in the reference's timers,
 * src/MolecularMechanics/hpc_minimization.cu:593). */
int stormm_device_synchronize(void);

/* ---- AtomGraphSynthesis ------------------------------------------------------------------- */
/* src/Synthesis/atomgraph_synthesis.h, constructor ag_synthesis_constructors.cpp:11-468 */
int stormm_synthesis_create(const stormm_topology_batch *batch, const stormm_launch_shape *shape,
                            stormm_synthesis **out);
int stormm_synthesis_destroy(stormm_synthesis *ag);
int stormm_synthesis_get_info(const stormm_synthesis *ag, stormm_synthesis_info *info);
int stormm_synthesis_atom_layout(const stormm_synthesis *ag, int32_t *atom_starts,
                                 int32_t *atom_counts);
/* Build the host-side plan (layout, work units, tile tasks) without touching the device. */
int stormm_synthesis_plan(const stormm_topology_batch *batch, const stormm_launch_shape *shape,
                          stormm_synthesis_info *info);
/* Kernel status bits (1 = coordinates beyond the LDS coordinate range); clear != 0 resets. */
int stormm_synthesis_status(const stormm_synthesis *ag, int32_t *status, int32_t clear);
/* Autotuner: time force + energy evaluations of the batch under candidate launch shapes (work-unit
 * capacity, units per CU, block size) on the device and return the fastest.  coordinates:
 * [sum atoms][3] A to time on (null: a loose lattice per system).  Replaces the reference's
 * launch-shape heuristics (src/Accelerator/core_kernel_manager.cpp:54-420). */
int stormm_autotune_launch_shape(const stormm_topology_batch *batch, int32_t precision,
                                 const double *coordinates, stormm_launch_shape *best,
                                 double *best_ms);

/* ---- PhaseSpaceSynthesis ------------------------------------------------------------------ */
/* src/Synthesis/phasespace_synthesis.h:325-383.  gpos_bits in [24, 72], frc_bits in [18, 72]
 * (src/Constants/fixed_precision.h:30-85); the overflow words are allocated with the arrays and
 * used above 38 coordinate bits and 40 force bits. */
int stormm_phasespace_create(const stormm_synthesis *ag, int32_t gpos_bits, int32_t frc_bits,
                             stormm_phasespace **out);
int stormm_phasespace_destroy(stormm_phasespace *ps);
/* xyz: [sum atoms][3] double (unpadded, system after system); host <-> device, synchronous */
int stormm_phasespace_upload_coordinates(stormm_phasespace *ps, const double *xyz);
int stormm_phasespace_download_coordinates(const stormm_phasespace *ps, double *xyz);
int stormm_phasespace_download_forces(const stormm_phasespace *ps, double *fxyz);
/* Raw fixed-precision copies, [sum atoms][3] int64 */
int stormm_phasespace_upload_fixed(stormm_phasespace *ps, const int64_t *xyz);
int stormm_phasespace_download_fixed(const stormm_phasespace *ps, int64_t *xyz, int64_t *fxyz);
/* Systems [lo, hi) only: PhaseSpaceSynthesis::upload / download(kind, lower, upper, gpu)
 * (src/Synthesis/phasespace_synthesis.h:649-672, .cpp:2327-2333,2379-2385).  xyz / fxyz hold
 * those systems' atoms, [atoms][3] double, system after system; fxyz may be NULL.  Invalid or
 * empty ranges return STORMM_ERR_INVALID_ARGUMENT.  Synchronous. */
int stormm_phasespace_upload_systems(stormm_phasespace *ps, int32_t lo, int32_t hi,
                                     const double *xyz);
int stormm_phasespace_download_systems(const stormm_phasespace *ps, int32_t lo, int32_t hi,
                                       double *xyz, double *fxyz);
/* One system: PhaseSpaceSynthesis::importSystem with its default TrajectoryKind::POSITIONS
 * (the raw-array and CoordinateFrame overloads, phasespace_synthesis.h:908-923, .tpp:9-) written
 * straight to the device: positions only.  The system's forces, prior positions and
 * conjugate-gradient history keep whatever they held, as in the reference, where forces and
 * velocities are imported only by the PhaseSpace overloads (.cpp:3410-3436); an evaluation
 * initialises forces itself and a minimization restarts its line search and gradient history.
 * exportSystem (phasespace_synthesis.h:698, .cpp:2472-) reads straight from the device
 * (coordinates and forces; xyz or fxyz may be NULL). */
int stormm_phasespace_import_system(stormm_phasespace *ps, int32_t system_index,
                                    const double *xyz);
int stormm_phasespace_export_system(const stormm_phasespace *ps, int32_t system_index,
                                    double *xyz, double *fxyz);
/* Raw fixed-precision values split into words and overflow words ([sum atoms][3] each; the
 * overflow arrays may be NULL, then every value must fit its 64-bit word). */
int stormm_phasespace_upload_fixed_split(stormm_phasespace *ps, const int64_t *xyz,
                                         const int32_t *xyz_ovrf);
int stormm_phasespace_download_fixed_split(const stormm_phasespace *ps, int64_t *xyz,
                                           int32_t *xyz_ovrf, int64_t *fxyz, int32_t *fxyz_ovrf);
/* The handle's device arrays in the PsSynthesisWriter layout (PhaseSpaceSynthesis::data(),
 * phasespace_synthesis.h:564). */
int stormm_phasespace_writer(stormm_phasespace *ps, stormm_ps_writer *w);
int stormm_phasespace_initialize_forces(stormm_phasespace *ps, void *stream);
int stormm_phasespace_prime_conjugate_gradient(stormm_phasespace *ps, void *stream);

/* ---- ScoreCard ---------------------------------------------------------------------------- */
/* src/Potential/scorecard.h:101-550 */
int stormm_scorecard_create(int32_t system_count, int32_t nrg_bits, stormm_scorecard **out);
int stormm_scorecard_destroy(stormm_scorecard *sc);
int stormm_scorecard_initialize(stormm_scorecard *sc, void *stream);
int stormm_scorecard_commit(stormm_scorecard *sc, void *stream);
/* instantaneous accumulators, [nsys][32] int64 */
int stormm_scorecard_download(const stormm_scorecard *sc, int64_t *accumulators);
/* running sums and sums of squares of committed samples, [nsys][32] double each */
int stormm_scorecard_download_history(const stormm_scorecard *sc, double *running,
                                      double *squared, int32_t *sample_count);
/* The time series (ScoreCard::time_series_accumulators, scorecard.h:503; reportHistory,
 * scorecard.cpp:1084-1134).  stormm_scorecard_commit is ScoreCard::commit followed by
 * incrementSampleCount (scorecard.cpp:366-, 608): it appends one sample, every system's row
 * with the potential energy sum in slot 28 and the total in slot 29, doubling the series'
 * capacity when it is full.  reserve sizes the series for `capacity` samples (ScoreCard::reserve,
 * scorecard.cpp:236-250); reset_sample_count keeps the first `count` samples (0 <= count <=
 * samples; resetSampleCount, :613) and leaves the running sums alone; set_last_time_step labels
 * the last sample (setLastTimeStep, :1142).  download_series copies samples [first, first +
 * count) as [count][nsys][32] int64 at nrg_bits, and their time steps when time_steps is not
 * NULL.  sample_count reports the samples held and the capacity. */
int stormm_scorecard_reserve(stormm_scorecard *sc, int32_t capacity);
int stormm_scorecard_reset_sample_count(stormm_scorecard *sc, int32_t count);
int stormm_scorecard_set_last_time_step(stormm_scorecard *sc, int32_t time_step);
int stormm_scorecard_sample_count(const stormm_scorecard *sc, int32_t *samples,
                                  int32_t *capacity);
int stormm_scorecard_download_series(const stormm_scorecard *sc, int32_t first, int32_t count,
                                     int64_t *series, int32_t *time_steps);
int64_t *stormm_scorecard_device_accumulators(stormm_scorecard *sc);
/* The handle's accumulators and history in the ScoreCardWriter layout (ScoreCard::data(),
 * scorecard.h:290-296); the time-series pointer holds until a commit or reserve grows the
 * series. */
int stormm_scorecard_data(stormm_scorecard *sc, stormm_scorecard_writer *w);

/* ---- MolecularMechanicsControls ----------------------------------------------------------- */
/* src/MolecularMechanics/mm_controls.h:100-363 */
int stormm_mm_controls_create(const stormm_minimize_controls *mincon,
                              stormm_mm_controls **out);
int stormm_mm_controls_destroy(stormm_mm_controls *ctrl);
int stormm_mm_controls_get_step(const stormm_mm_controls *ctrl, int32_t *step);
int stormm_mm_controls_set_step(stormm_mm_controls *ctrl, int32_t step);
/* MolecularMechanicsControls::dpData / spData (mm_controls.cpp:300-320). */
int stormm_mm_controls_data(const stormm_mm_controls *ctrl, stormm_mm_kit *kit);

/* ---- LineMinimization --------------------------------------------------------------------- */
/* src/MolecularMechanics/line_minimization.h:85-221 */
int stormm_line_min_create(int32_t system_count, double dx0, stormm_line_min **out);
int stormm_line_min_destroy(stormm_line_min *lm);
/* [nsys] each: l_move, s_move, mfac_a, mfac_b, mfac_c, nrg_a, nrg_b, nrg_c, nrg_d */
int stormm_line_min_download(const stormm_line_min *lm, double *values9xN);
/* LineMinimization::data (line_minimization.h:160-166). */
int stormm_line_min_data(stormm_line_min *lm, stormm_linmin_writer *w);

/* ---- Launch entry points ------------------------------------------------------------------ */
/* Coulomb + Lennard-Jones + (if the synthesis has an implicit-solvent model) Born radii, GB pair
 * energies and the radius-derivative pass.  Forces are added into the PhaseSpaceSynthesis force
 * arrays and energies into the ScoreCard instantaneous accumulators.  clash_minimum_distance
 * and clash_ratio are the reference's trailing arguments (hpc_nonbonded_potential.h:131-157):
 * when either exceeds 10^-6, Coulomb is softened inside that distance and Lennard-Jones inside
 * that multiple of sigma_ij (pass 0, 0 for the plain potentials). */
int stormm_launch_nonbonded(int32_t precision, const stormm_synthesis *ag,
                            stormm_mm_controls *ctrl, stormm_phasespace *ps,
                            stormm_scorecard *sc, int32_t eval_force, int32_t eval_energy,
                            double clash_minimum_distance, double clash_ratio, void *stream);
/* Bonds, angles, cosine dihedrals and attenuated 1:4 pairs (clash damping of the 1:4 pairs as
 * in hpc_valence_potential.h:356-378). */
int stormm_launch_valence(int32_t precision, const stormm_synthesis *ag,
                          stormm_mm_controls *ctrl, stormm_phasespace *ps, stormm_scorecard *sc,
                          int32_t eval_force, int32_t eval_energy, double clash_minimum_distance,
                          double clash_ratio, void *stream);
/* All terms in one fused launch. */
int stormm_launch_mm_evaluation(int32_t precision, const stormm_synthesis *ag,
                                stormm_mm_controls *ctrl, stormm_phasespace *ps,
                                stormm_scorecard *sc, int32_t eval_force, int32_t eval_energy,
                                double clash_minimum_distance, double clash_ratio,
                                void *stream);
/* Polak-Ribiere direction mixing and normalization of the force arrays per system. */
int stormm_launch_conjugate_gradient(int32_t precision, const stormm_synthesis *ag,
                                     stormm_phasespace *ps, stormm_mm_controls *ctrl,
                                     void *stream);
/* One of the four line moves (move_number 0..3). */
int stormm_launch_line_advance(int32_t precision, const stormm_synthesis *ag,
                               stormm_phasespace *ps, stormm_scorecard *sc, stormm_line_min *lm,
                               int32_t move_number, void *stream);
/* Whole minimization.  persistent == 0 issues the reference's per-step launch sequence (per
 * cycle: force evaluation, conjugate gradient, four line moves with three energy evaluations);
 * persistent == 2 captures that sequence for one cycle in a HIP graph and replays it for every
 * cycle.  persistent == 1, 3 and 4 run fused cycle kernels: each evaluation block also updates
 * the conjugate-gradient direction and makes the line move of the whole systems it holds, so a
 * cycle is four launches; mode 1 captures the whole run in one HIP graph and launches it once,
 * mode 3 captures one cycle and replays it, mode 4 issues the kernels one by one.  All modes
 * give bit-identical results.  Modes 1-4 stop a system once its RMS force reaches
 * convergence_rms_force (if > 0) and record per-system cycle counts and RMS forces
 * (stormm_minimization_status).  Final energies land in the ScoreCard.  Modes 1, 3 and 4 take
 * no restraints and at most 38 coordinate and 40 force bits (STORMM_ERR_STATE otherwise); a
 * restrained synthesis is minimized in modes 0 and 2.  Mode 0 returns once its launches are
 * queued on the stream; the other modes return when the minimization is done.
 * With diagnostic_print_frequency = ntpr > 0 every mode commits the ScoreCard after the force
 * evaluation of each cycle that is a multiple of ntpr (hpc_minimization.cu:292-300): the
 * energies at the start of that cycle, appended to the time series and the running sums, with
 * the cycle as the sample's time step; ceil(total_cycles / ntpr) samples per call.  A system
 * that stops early repeats its last sample in the ones it did not reach.  The kit forms write
 * the writer's time_series from sampled_step_count on (the caller advances its own count; a
 * series too small returns STORMM_ERR_CAPACITY; none given: no samples). */
int stormm_launch_minimization(int32_t precision, const stormm_synthesis *ag,
                               stormm_phasespace *ps, const stormm_minimize_controls *mincon,
                               stormm_scorecard *sc, stormm_line_min *lm, int32_t persistent,
                               void *stream);
/* Per-system status after a minimization in modes 1-4: cycles run and final RMS force. */
int stormm_minimization_status(const stormm_synthesis *ag, int32_t *cycles_run,
                               double *rms_force);

/* ---- Launch entry points on caller-owned arrays (kit form) --------------------------------- */
/* The same launches as above with the state in the caller's device arrays, passed in the
 * reference's abstract layouts, as its launch functions take them (launchNonbonded
 * hpc_nonbonded_potential.h:124-140, launchValence hpc_valence_potential.h:167-210,
 * launchConjugateGradient hpc_reduction.h:55-63, launchLineAdvance hpc_minimization.h:63-70,
 * launchMinimization hpc_minimization.h:122-151).  The parameters still come from the synthesis
 * handle; the arrays must follow its atom layout (stormm_synthesis_atom_layout, checked by
 * stormm_check_ps_writer).  A NULL controls kit means step 0.  Forces above 40 bits are
 * accumulated at 40 bits within a work unit and stored at frc_bits; single-precision launches
 * take at most 40 force bits.  Mode 1 (persistent) of the minimization takes no overflow words
 * (gpos_bits <= 38, frc_bits <= 40). */
int stormm_check_ps_writer(const stormm_synthesis *ag, const stormm_ps_writer *psw);
int stormm_launch_nonbonded_kit(int32_t precision, const stormm_synthesis *ag,
                                const stormm_mm_kit *ctrl, const stormm_ps_writer *psw,
                                const stormm_scorecard_writer *scw, int32_t eval_force,
                                int32_t eval_energy, double clash_minimum_distance,
                                double clash_ratio, void *stream);
int stormm_launch_valence_kit(int32_t precision, const stormm_synthesis *ag,
                              const stormm_mm_kit *ctrl, const stormm_ps_writer *psw,
                              const stormm_scorecard_writer *scw, int32_t eval_force,
                              int32_t eval_energy, double clash_minimum_distance,
                              double clash_ratio, void *stream);
int stormm_launch_mm_evaluation_kit(int32_t precision, const stormm_synthesis *ag,
                                    const stormm_mm_kit *ctrl, const stormm_ps_writer *psw,
                                    const stormm_scorecard_writer *scw, int32_t eval_force,
                                    int32_t eval_energy, double clash_minimum_distance,
                                    double clash_ratio, void *stream);
int stormm_launch_conjugate_gradient_kit(int32_t precision, const stormm_synthesis *ag,
                                         const stormm_ps_writer *psw, const stormm_mm_kit *ctrl,
                                         void *stream);
int stormm_launch_line_advance_kit(int32_t precision, const stormm_synthesis *ag,
                                   const stormm_ps_writer *psw,
                                   const stormm_scorecard_writer *scw,
                                   const stormm_linmin_writer *lmw, int32_t move_number,
                                   void *stream);
int stormm_launch_minimization_kit(int32_t precision, const stormm_synthesis *ag,
                                   const stormm_ps_writer *psw,
                                   const stormm_minimize_controls *mincon,
                                   const stormm_scorecard_writer *scw,
                                   const stormm_linmin_writer *lmw, int32_t persistent,
                                   void *stream);
/* Allocation-free minimization on caller-owned arrays.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under stream capture, tight serving
 * loops).  Runs mode 1's fused-cycle kernels, queued one by one on `stream` without a graph and
 * without waiting (the call returns once the launches are queued), with each system's cycle
 * count, stop flag and RMS force in a static device array of 1 MiB per device that the library
 * holds until the process ends: at most 65536 systems (more returns STORMM_ERR_CAPACITY).  The
 * array starts every call cleared; calls that overlap in time on one device share it and race.
 * Mode 1's limits apply: no restraints, at most 38 coordinate and 40 force bits
 * (STORMM_ERR_STATE otherwise).  Results equal stormm_launch_minimization_kit in mode 1. */
int stormm_launch_minimization_kit_noalloc(int32_t precision, const stormm_synthesis *ag,
                                           const stormm_ps_writer *psw,
                                           const stormm_minimize_controls *mincon,
                                           const stormm_scorecard_writer *scw,
                                           const stormm_linmin_writer *lmw, void *stream);
/* PhaseSpaceSynthesis::initializeForces / primeConjugateGradientCalculation and
 * ScoreCard::initialize on caller-owned arrays. */
int stormm_ps_writer_initialize_forces(const stormm_ps_writer *psw, int32_t padded_atoms,
                                       void *stream);
int stormm_ps_writer_prime_conjugate_gradient(const stormm_ps_writer *psw, int32_t padded_atoms,
                                              void *stream);
int stormm_scorecard_writer_initialize(const stormm_scorecard_writer *scw, void *stream);

/* ---- Timing helpers (HIP events on the launch stream) -------------------------------------- */
/* Run `repeats` fused evaluations back to back on `stream` and return the mean kernel time in
 * milliseconds measured with HIP events recorded on that stream. */
int stormm_time_mm_evaluation(int32_t precision, const stormm_synthesis *ag,
                              stormm_phasespace *ps, stormm_scorecard *sc, int32_t eval_force,
                              int32_t repeats, void *stream, double *mean_ms);
/* The same for one choice of terms (the launches of stormm_launch_nonbonded,
 * stormm_launch_valence or both fused): the benchmark's breakdown of where the time goes. */
#define STORMM_TERMS_NONBONDED 1
#define STORMM_TERMS_VALENCE 2
#define STORMM_TERMS_ALL 3
int stormm_time_launch(int32_t precision, const stormm_synthesis *ag, stormm_phasespace *ps,
                       stormm_scorecard *sc, int32_t terms, int32_t eval_force,
                       int32_t repeats, void *stream, double *mean_ms);

/* ---- Serving pool ------------------------------------------------------------------------- */
/* No counterpart in the reference, which minimizes one fixed batch per launchMinimization call
 * (src/MolecularMechanics/hpc_minimization.cu:129-596).  A pool keeps systems resident on the
 * device with all of their minimization state and advances whatever is resident a few cycles at
 * a time, so that requests can join and leave between calls (continuous batching).  Each
 * admission takes device memory from a fixed budget; a system stops at the cycle limit or when
 * its RMS force reaches the threshold.  A system's trajectory does not depend on what else is
 * resident or on how its cycles are split across calls. */
typedef struct stormm_pool stormm_pool;
typedef struct {
  int32_t precision;               /* STORMM_PREC_SINGLE or STORMM_PREC_DOUBLE                  */
  int32_t gpos_bits, frc_bits, nrg_bits;
  int64_t hbm_budget_bytes;        /* device memory for resident systems and their templates     */
  int32_t max_systems;             /* resident system slots                                     */
  int32_t igb;                     /* implicit-solvent model shared by every admitted system     */
  double dielectric;
  double salt_concentration;
  double coulomb_constant;
  stormm_minimize_controls minimize; /* cycle limit, steepest-descent cycles, dx0, RMS stop    */
  stormm_launch_shape shape;       /* work-unit shaping (zero = defaults)                       */
} stormm_pool_config;

int stormm_pool_create(const stormm_pool_config *config, stormm_pool **out);
int stormm_pool_destroy(stormm_pool *pool);
/* Free atoms and slots, bytes held by work templates, and the atom capacity of the budget. */
int stormm_pool_capacity(const stormm_pool *pool, int64_t *free_atoms, int32_t *free_slots,
                         int64_t *template_bytes, int64_t *capacity_atoms);
/* Admit every system of `batch` with coordinates xyz ([sum atoms][3] A); writes their slots.
 * STORMM_ERR_CAPACITY (nothing admitted) when the budget or the slots cannot hold them. */
int stormm_pool_admit(stormm_pool *pool, const stormm_topology_batch *batch, const double *xyz,
                      int32_t *slots);
/* Admit `count` conformers of the one system in `molecule`; xyz is [count][N][3] A. */
int stormm_pool_admit_conformers(stormm_pool *pool, const stormm_topology_batch *molecule,
                                 int32_t count, const double *xyz, int32_t *slots);
/* Return systems' slots and (once all systems of an admission have left) their atoms. */
int stormm_pool_release(stormm_pool *pool, const int32_t *slots, int32_t count);
/* Advance every resident, still-running system by up to `cycles` minimization cycles (the
 * reference's per-cycle launch sequence) and evaluate its energy at the coordinates reached;
 * writes how many systems were running.  Work is queued on `stream`. */
int stormm_pool_minimize(stormm_pool *pool, int32_t cycles, int32_t *running, void *stream);
/* Per slot: cycles completed, finished flag (threshold or cycle limit), last RMS force
 * (kcal/mol/A) and potential energy (kcal/mol) at the current coordinates. */
int stormm_pool_status(const stormm_pool *pool, const int32_t *slots, int32_t count,
                       int32_t *steps, int32_t *done, double *rms, double *energy);
/* Energy history of one slot, the per-system form of the ScoreCard time series (ScoreCard::
 * reportHistory, src/Potential/scorecard.cpp:1090-1102): with minimize.
 * diagnostic_print_frequency = ntpr > 0 the pool keeps ceil(total_cycles / ntpr) samples per
 * slot, sample k being the energies at the start of the slot's cycle k * ntpr counted from its
 * admission (plus the cycle at which the RMS force stopped it).  Writes the samples recorded so
 * far to count and, for each, the potential energy (kcal/mol) to energy and the ScoreCard row
 * (32 int64 at nrg_bits, potential energy in slot 28) to rows; either may be NULL.  More
 * samples than `capacity` returns STORMM_ERR_CAPACITY; no history configured gives count 0. */
int stormm_pool_history(const stormm_pool *pool, int32_t slot, int32_t capacity, int32_t *count,
                        double *energy, int64_t *rows);
/* Coordinates of one system, [N][3] in A and/or fixed precision (either pointer may be null). */
int stormm_pool_coordinates(const stormm_pool *pool, int32_t slot, double *xyz,
                            int64_t *fixed_xyz);

/* ---- Device microbenchmarks (roofline calibration) ------------------------------------------ */
/* Stream-copy `bytes` bytes `repeats` times; returns read+write GB/s.  No reference counterpart:
 * used to confirm the HBM ceiling on the device and to scale profiler memory counters. */
int stormm_copy_bandwidth(int64_t bytes, int32_t repeats, double *gbps);
/* Achievable non-packed FP32 FMA rate in TFLOP/s.  No reference counterpart. */
int stormm_fma_throughput(int32_t iters, double *tflops);

#ifdef __cplusplus
}
#endif
#endif
