#!/usr/bin/env python3
"""Throughput benchmark of the batched energy + force evaluation on one MI355X.

Workload: 1024 heterogeneous synthetic branched molecules with atom counts ~ U[24, 256] (seeded),
OBC-II Generalized Born in a dielectric of 78.5, Coulomb + Lennard-Jones over all non-excluded
pairs, bonds / angles / cosine dihedrals / attenuated 1:4 pairs; single-precision pair arithmetic
with 23-bit fixed-precision force accumulation.  One step = one batched energy + force evaluation
of all 1024 systems (force and ScoreCard zeroing, then the fused evaluation launch), with
coordinates and parameters resident in device memory.

Prints one JSON line.  The double-precision mode is timed alongside, and both modes' errors
against the FP64 CPU oracle are measured on a 16-system subsample.  The CPU baseline times the
oracle (the restatement of the reference's CPU path) on a bounded sample of the same systems with
one thread per core of the job's CPU share.

Further workloads ride on the same line (skip them with --no-extra):
  "lj_table": the same 1024 molecules with 16 Lennard-Jones types each and stored T x T tables
    (one pair of types off the combining rule, as an NBFIX pair is), the path an AMBER topology
    takes: kernel time next to the combining-rule path's, and FP32 errors on 16 systems.
  "minimization": 4096 systems, N ~ U[16, 128], starts perturbed by N(0, 0.3 A), 500 cycles
    (50 steepest descent, then conjugate gradient, the first 25 clash-damped) in FP32: the fused
    cycle kernels captured whole in one HIP graph (whole batch, and once per size class), the
    reference's per-step launch sequence, and that sequence captured per cycle and replayed,
    with the oracle's minimizer on host threads as the CPU rate.
  "large_batch": 16384 systems, N ~ U[32, 96]: FP32 and FP64 evaluation throughput and the
    errors of both modes against the FP64 oracle on a 64-system subsample.
  "serving": the served-minimization stream (Poisson arrivals at 200 requests/s over
    --serve-seconds, default 60, 1-64 conformers of one molecule with N ~ logU[16, 512] per
    request, 300 cycles or RMS force <= 0.1): continuous batching and static batching in real
    time (p50 / p99 latency), and the capacity of continuous serving with every request waiting
    against static serving, against the batch minimizer given all of the same conformers in one
    launch, and against one batched launch per request.

  python bench.py [--gpus 1] [--steps 500] [--warmup 50] [--cpu-seconds 12] [--no-cpu]
                  [--no-extra]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("system energy+force evals/s (valence+NB+GB); served minimizations/s, p99 latency")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads():
    """Host threads for the CPU baseline: one per core of this job's CPU share.  On the GPU
    machine a one-GPU job is given 16 cores (OMP_NUM_THREADS carries the share); the
    affinity mask may list more cores that other jobs hold."""
    visible = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    if share and share.isdigit() and int(share) > 0:
        return min(visible, int(share)), visible
    return visible, visible


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def size_stratified(natoms, count):
    """`count` system indices spread evenly over the batch sorted by N^2 (pair work), so the
    sample's pair-work distribution follows the batch's."""
    order = np.argsort(np.asarray(natoms) ** 2, kind="stable")
    picks = np.linspace(0, len(order) - 1, count).round().astype(int)
    return [int(order[i]) for i in picks]


def parity_errors(S, oracle, topos, coords, prec, gpos, frc, nrg):
    ag = S.AtomGraphSynthesis(topos)
    ps = S.PhaseSpaceSynthesis(ag, gpos, frc)
    sc = S.ScoreCard(ag.system_count, nrg)
    ps.upload(coords)
    ps.initialize_forces()
    sc.initialize()
    S.launch_mm_evaluation(prec, ag, None, ps, sc, True, True)
    E = sc.accumulators()[:, :13] * 2.0 ** -nrg
    F = ps.forces()
    X = ps.coordinates()
    worst_e, worst_et, worst_f = 0.0, 0.0, 0.0
    # Per energy component, relative to that component (components of at least 1 kcal/mol).
    worst_c = {}
    for s, t in enumerate(topos):
        eo_d, _, fo = oracle.eval_mm(t, X[s], True, nrg)
        eo = np.array([eo_d[k] for k in oracle.SLOT_NAMES])
        scale = np.abs(eo).sum()
        worst_e = max(worst_e, abs(E[s].sum() - eo.sum()) / scale)
        worst_et = max(worst_et, abs(E[s].sum() - eo.sum()) / abs(eo.sum()))
        worst_f = max(worst_f, float(np.sqrt(((F[s] - fo) ** 2).sum() / (fo ** 2).sum())))
        for k, name in enumerate(oracle.SLOT_NAMES):
            if abs(eo[k]) >= 1.0:
                worst_c[name] = max(worst_c.get(name, 0.0), abs(E[s][k] - eo[k]) / abs(eo[k]))
    return {"energy_rel_sum_abs_terms": worst_e, "energy_rel_total": worst_et,
            "energy_rel_component": worst_c, "force_rel_rms": worst_f}


def minimization_workload(S, synthetic, oracle, nsys, cycles, with_cpu, shape=None):
    """FP32 minimization of the 4096-system batch: three launch strategies and the CPU rate."""
    topos, coords = synthetic.config2(nsys=nsys, seed=2)
    sd, dx0 = 50, 0.01
    mc = S.minimize_controls(cycles, sd, dx0, 0.0)
    if shape is None:
        # Work-unit shape from the autotuner (timed force evaluations of the whole batch).
        shape, tuned_ms = S.autotune(topos, "single", coords)

    def setup(sub_t, sub_c, sub_shape=None):
        ag = S.AtomGraphSynthesis(sub_t, sub_shape or shape)
        ps = S.PhaseSpaceSynthesis(ag, 32, 23)
        sc = S.ScoreCard(ag.system_count, 26)
        lm = S.LineMinimization(ag.system_count, dx0)
        ps.upload(sub_c)
        return ag, ps, sc, lm

    def run(parts, persistent):
        S.synchronize()
        t0 = time.perf_counter()
        for ag, ps, sc, lm in parts:
            S.launch_minimization("single", ag, ps, mc, sc, lm, persistent=persistent)
        S.synchronize()
        return time.perf_counter() - t0

    # Warm-up on a few systems (kernel loading).
    wag, wps, wsc, wlm = setup(topos[:8], coords[:8])
    for mode in (1, 3, 4, 0, 2):
        S.launch_minimization("single", wag, wps, S.minimize_controls(3, 1, dx0, 0.0), wsc, wlm,
                              persistent=mode)
    # Starting energies of every system (one energy evaluation at the perturbed coordinates).
    ag0, ps0, sc0, _ = setup(topos, coords)
    sc0.initialize()
    S.launch_mm_evaluation("single", ag0, None, ps0, sc0, False, True)
    e_start = sc0.potential_energies()
    del ag0, ps0, sc0
    timings, energies = {}, {}
    for mode, name in ((1, "fused_graph"), (3, "fused_cycle_graph"), (4, "fused_direct"),
                       (0, "per_step"), (2, "graph_replay")):
        whole = [setup(topos, coords)]
        timings[name] = run(whole, mode)
        energies[name] = whole[0][2].potential_energies()
        del whole
    # The fused whole-run graph once per size class (atom-count quartiles of the range).
    edges = [16, 44, 72, 100, 129]
    classes = [[i for i, t in enumerate(topos) if lo <= t.natom < hi]
               for lo, hi in zip(edges[:-1], edges[1:])]
    classes = [idx for idx in classes if idx]
    parts = [setup([topos[i] for i in idx], [coords[i] for i in idx]) for idx in classes]
    t_class = run(parts, 1)
    # Every system ends where it does in the whole-batch launch (results do not depend on the
    # batch's composition).
    e_class = np.empty(nsys)
    for idx, (_, _, sc_c, _) in zip(classes, parts):
        e_class[idx] = sc_c.potential_energies()
    del parts
    best = min(timings, key=timings.get)
    out = {"systems": nsys, "cycles": cycles, "sd_cycles": sd, "precision": "fp32",
           "clash_damping_cycles": 25, "autotuned_shape": shape,
           "fused_graph_s": timings["fused_graph"],
           "fused_cycle_graph_s": timings["fused_cycle_graph"],
           "fused_direct_s": timings["fused_direct"],
           "per_step_launches_s": timings["per_step"],
           "graph_replay_s": timings["graph_replay"],
           "per_size_class_s": t_class, "size_classes": len(classes),
           "minimizations_per_s": nsys / timings[best],
           "best_mode": best,
           "fused_vs_per_step": timings["per_step"] / min(timings["fused_graph"],
                                                          timings["fused_cycle_graph"],
                                                          timings["fused_direct"]),
           "fused_graph_vs_per_step": timings["per_step"] / timings["fused_graph"],
           "graph_vs_per_step": timings["per_step"] / timings["graph_replay"],
           "fused_equals_per_step": bool(all(np.array_equal(energies[m], energies["per_step"])
                                             for m in ("fused_graph", "fused_cycle_graph",
                                                       "fused_direct"))),
           "graph_equals_per_step": bool(np.array_equal(energies["graph_replay"],
                                                        energies["per_step"])),
           "per_size_class_equals_whole_batch": bool(np.array_equal(e_class,
                                                                    energies["fused_graph"])),
           "every_energy_fell": bool(np.all(energies["fused_graph"] < e_start)),
           "mean_energy_start": float(np.mean(e_start)),
           "mean_energy_final": float(np.mean(energies["fused_graph"]))}
    if with_cpu:
        from concurrent.futures import ThreadPoolExecutor
        nsample = 128
        threads, visible = cpu_threads()
        pick = size_stratified([t.natom for t in topos], nsample)
        fixed = {i: np.rint(np.asarray(coords[i]) * 2.0 ** 32).astype(np.int64) for i in pick}

        def one(i):
            oracle.minimize(topos[i], fixed[i], 32, 23, 26, cycles, sd, dx0, dp_mode=False,
                            clash_cycles=min(25, cycles))

        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(one, pick))
        secs = time.perf_counter() - t0
        # The sample's share of the batch's pair work scales its rate to the whole batch.
        n2 = np.array([t.natom for t in topos], dtype=np.float64) ** 2
        work_share = n2[pick].sum() / n2.sum()
        out["cpu_baseline"] = {"value": nsys / (secs / work_share),
                               "unit": "minimizations/s", "cores": threads, "kind": "port",
                               "sample": f"{nsample} of the {nsys} systems spread evenly over "
                                         f"the batch sorted by N^2 ({100 * work_share:.2f} % of "
                                         f"its pair work), {cycles} cycles each, {threads} "
                                         f"threads, {secs:.1f} s; rate = batch systems / "
                                         f"(time scaled by pair work)",
                               "host_cores_visible": visible, "cpu_model": cpu_model(),
                               "compiler": "gcc " + " ".join(oracle.CFLAGS)}
    return out


def stage_breakdown(S, topos, coords, reps):
    """Where the FP32 evaluation's time goes, from launches of the library's own entry points on
    the headline batch: the fused evaluation with and without forces, the non-bonded launch
    (Born integrals, radii, Coulomb + Lennard-Jones + GB pairs, radius derivatives) with and
    without forces, the valence launch, and the non-bonded launch of the same molecules in
    vacuum (Coulomb + Lennard-Jones pairs only)."""
    import copy
    out = {}

    def timed(tops, runs):
        ag = S.AtomGraphSynthesis(tops)
        ps = S.PhaseSpaceSynthesis(ag, 32, 23)
        sc = S.ScoreCard(ag.system_count, 26)
        ps.upload(coords)
        for name, terms, force in runs:
            S.time_launch("single", ag, ps, sc, terms, force, 5)
            out[name] = S.time_launch("single", ag, ps, sc, terms, force, reps)

    timed(topos, [("fused_force_ms", "all", True), ("fused_energy_ms", "all", False),
                  ("nonbonded_force_ms", "nonbonded", True),
                  ("nonbonded_energy_ms", "nonbonded", False),
                  ("valence_force_ms", "valence", True)])
    vacuum = []
    for t in topos:
        v = copy.copy(t)
        v.igb = 0
        vacuum.append(v)
    timed(vacuum, [("vacuum_nonbonded_force_ms", "nonbonded", True)])
    out["gb_derivative_and_pair_forces_ms"] = (out["nonbonded_force_ms"] -
                                               out["nonbonded_energy_ms"])
    out["note"] = ("kernel times (HIP events, mean of the repeats) of separate launches; the "
                   "fused evaluation overlaps valence with the Born-integral tiles, so the parts "
                   "do not add up to it exactly")
    return out


def table_topologies(topos, ntypes=16):
    """The same molecules with AMBER-style Lennard-Jones type tables: atom i takes type i mod
    ntypes (the type's sigma and epsilon are those of its first atom), the T x T A / B tables are
    built by Lorentz-Berthelot combining, and one pair of types departs from it by 5 % (as an
    NBFIX pair does), so every system keeps its stored tables."""
    import copy
    out = []
    for t in topos:
        types = (np.arange(t.natom) % ntypes).astype(np.int32)
        first = [int(np.flatnonzero(types == k)[0]) for k in range(ntypes)]
        sig = np.asarray(t.lj_sigma, dtype=np.float64)[first]
        eps = np.asarray(t.lj_epsilon, dtype=np.float64)[first]
        sij = 0.5 * (sig[:, None] + sig[None, :])
        eij = np.sqrt(eps[:, None] * eps[None, :])
        a = 4.0 * eij * sij ** 12
        b = 4.0 * eij * sij ** 6
        a[0, 1] *= 1.05
        a[1, 0] = a[0, 1]
        v = copy.copy(t)
        v.lj_idx, v.n_lj_types = types, ntypes
        v.lj_sigma = v.lj_epsilon = None
        v.lja, v.ljb = a.reshape(-1), b.reshape(-1)
        v.lja14, v.ljb14 = a.reshape(-1).copy(), b.reshape(-1).copy()
        out.append(v)
    return out


def lj_table_workload(S, oracle, topos, coords, reps):
    """The headline batch on the stored-table Lennard-Jones path: kernel time against the
    combining-rule path of the same molecules, and oracle errors on 16 systems."""
    tab = table_topologies(topos)
    ag = S.AtomGraphSynthesis(tab)
    ps = S.PhaseSpaceSynthesis(ag, 32, 23)
    sc = S.ScoreCard(ag.system_count, 26)
    ps.upload(coords)
    S.time_mm_evaluation("single", ag, ps, sc, True, 10)
    ms = S.time_mm_evaluation("single", ag, ps, sc, True, reps)
    out = {"systems": len(tab), "lj_types_per_system": 16,
           "table_systems": int(ag.info.lj_table_systems), "kernel_ms": ms,
           "system_evals_per_s": len(tab) / (ms * 1.0e-3)}
    out["errors_fp32"] = parity_errors(S, oracle, tab[:16], coords[:16], "single", 32, 23, 26)
    return out


def large_batch_workload(S, synthetic, oracle, nsys, reps):
    """16384-system evaluation throughput in both precision modes and their oracle errors."""
    topos, coords = synthetic.config4(nsys=nsys, seed=4)
    from stormm_amd import roofline
    out = {"systems": nsys}
    for prec, bits in (("single", (32, 23, 26)), ("double", (32, 40, 32))):
        # Smaller molecules than the headline batch: each precision on its own tuned shape.
        shape, _ = S.autotune(topos, prec, coords)
        ag = S.AtomGraphSynthesis(topos, shape)
        ps = S.PhaseSpaceSynthesis(ag, bits[0], bits[1])
        sc = S.ScoreCard(ag.system_count, bits[2])
        ps.upload(coords)
        S.time_mm_evaluation(prec, ag, ps, sc, True, 3)
        ms = S.time_mm_evaluation(prec, ag, ps, sc, True, reps)
        info = ag.info
        flops = roofline.eval_flops(info.pair_count, info.valence_term_count, ag.total_atoms)
        nbytes = roofline.eval_bytes(ag.total_atoms, info.valence_term_count, info.task_count)
        peak = roofline.PEAK_FP32_TFLOPS if prec == "single" else roofline.PEAK_FP64_TFLOPS
        tflops = flops / (ms * 1.0e-3) / 1.0e12
        out["fp32" if prec == "single" else "fp64"] = {
            "kernel_ms": ms, "system_evals_per_s": nsys / (ms * 1.0e-3),
            "autotuned_shape": shape,
            "force_bits": bits[1], "energy_bits": bits[2],
            "pairs": info.pair_count, "flop_per_launch": flops,
            "achieved_tflops": tflops, "valu_peak_tflops": peak, "valu_frac": tflops / peak,
            "algorithmic_bytes_per_launch": nbytes,
            "achieved_hbm_gbps": nbytes / (ms * 1.0e-3) / 1.0e9,
            "hbm_frac": nbytes / (ms * 1.0e-3) / 1.0e9 / roofline.PEAK_HBM_GBPS}
        del ag, ps, sc
    sub = slice(0, 64)
    out["errors"] = {
        "fp32": parity_errors(S, oracle, topos[sub], coords[sub], "single", 32, 23, 26),
        "fp64": parity_errors(S, oracle, topos[sub], coords[sub], "double", 32, 40, 32),
        "sample": "first 64 of the 16384 systems, oracle at the device coordinates"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-parity", action="store_true", help="skip the oracle error check")
    ap.add_argument("--systems", type=int, default=1024)
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the minimization and large-batch workloads")
    ap.add_argument("--min-cycles", type=int, default=500)
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="arrival window of the served-minimization stream")
    ap.add_argument("--minimize-only", metavar="CAPACITY,PER_CU,THREADS", default=None,
                    help="run only the minimization workload on the given work-unit shape (no "
                         "autotuning, no CPU baseline) and print its JSON block; for profiling")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this path runs every system of a batch on a single "
              f"device and has no multi-GPU mode (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    from stormm_amd import roofline, synthetic
    from stormm_amd import synthesis as S

    if args.minimize_only:
        cap, per_cu, threads = (int(v) for v in args.minimize_only.split(","))
        shape = {"wu_atom_capacity": cap, "target_wu_per_cu": per_cu, "block_threads": threads}
        res = minimization_workload(S, synthetic, None, 4096, args.min_cycles, False, shape)
        print(json.dumps({"minimization": res}), flush=True)
        return

    log(f"device: {S.device_info()}")
    t0 = time.time()
    topos, coords = synthetic.config1(nsys=args.systems, seed=1)
    log(f"generated {len(topos)} systems in {time.time() - t0:.1f} s")
    ag = S.AtomGraphSynthesis(topos)
    info = ag.info
    ps = S.PhaseSpaceSynthesis(ag, 32, 23)
    sc = S.ScoreCard(ag.system_count, 26)
    ps.upload(coords)
    log(f"work units {info.work_unit_count}, tasks {info.task_count}, pairs {info.pair_count}, "
        f"valence terms {info.valence_term_count}")

    def step(prec):
        ps.initialize_forces()
        sc.initialize()
        S.launch_mm_evaluation(prec, ag, None, ps, sc, True, True)

    for _ in range(args.warmup):
        step("single")
    S.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step("single")
    S.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1.0e3 * wall / args.steps
    value = info.system_count * args.steps / wall
    log(f"FP32: {ms_per_step:.4f} ms/step, {value:.4e} system-evals/s")

    # Kernel duration from HIP events on the launch stream.
    reps = max(50, args.steps)
    kernel_ms = S.time_mm_evaluation("single", ag, ps, sc, True, reps)
    # FP64 mode on a synthesis whose work-unit shape is tuned for it (fewer registers per lane
    # are free in double precision, so smaller blocks and more units per CU pay).
    shape64, _ = S.autotune(topos, "double", coords)
    ag64 = S.AtomGraphSynthesis(topos, shape64)
    ps64 = S.PhaseSpaceSynthesis(ag64, 32, 40)
    sc64 = S.ScoreCard(ag64.system_count, 32)
    ps64.upload(coords)
    S.time_mm_evaluation("double", ag64, ps64, sc64, True, 10)
    kernel_ms_fp64 = S.time_mm_evaluation("double", ag64, ps64, sc64, True, max(20, reps // 4))
    del ag64, ps64, sc64
    atoms = ag.total_atoms
    flops = roofline.eval_flops(info.pair_count, info.valence_term_count, atoms, True)
    achieved = flops / (kernel_ms * 1.0e-3) / 1.0e12
    bytes_alg = roofline.eval_bytes(atoms, info.valence_term_count, info.task_count, True)
    log(f"kernel {kernel_ms:.4f} ms (FP64 {kernel_ms_fp64:.4f} ms); {flops / 1e9:.3f} GFLOP "
        f"-> {achieved:.2f} TFLOP/s")
    breakdown = stage_breakdown(S, topos, coords, max(50, args.steps // 2))
    log(f"stage breakdown: {breakdown}")
    measured_fma = S.fma_throughput()
    measured_bw = S.copy_bandwidth(1 << 30, 10)
    log(f"measured ceilings: FP32 FMA {measured_fma:.1f} TFLOP/s, copy {measured_bw:.0f} GB/s")

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_mm_eval_fp32_1024sys_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("systems") == info.system_count and tr.get("pairs") == info.pair_count:
            traffic = tr.get("hbm_bytes_per_launch")

    parity = None
    if not args.no_parity:
        import oracle
        sub_t, sub_c = topos[:16], coords[:16]
        parity = {"fp32": parity_errors(S, oracle, sub_t, sub_c, "single", 32, 23, 26),
                  "fp64": parity_errors(S, oracle, sub_t, sub_c, "double", 32, 40, 32),
                  "sample": "first 16 of the 1024 systems, oracle at the device coordinates"}
        log(f"parity: {parity}")

    cpu = None
    if not args.no_cpu:
        import oracle
        threads, visible = cpu_threads()
        # The whole batch, one thread per core of the job's share, systems split into
        # contiguous ranges of equal pair work.
        batch = oracle.Batch(topos, coords)
        secs, _ = batch.run(repeats=1, nthreads=threads)
        reps_cpu = int(max(1, min(5000, round(args.cpu_seconds / max(secs, 1e-6)))))
        secs, _ = batch.run(repeats=reps_cpu, nthreads=threads)
        cpu_value = len(topos) * reps_cpu / secs
        # Single-core rate on a pair-work-stratified quarter of the batch.
        pick = size_stratified([t.natom for t in topos], len(topos) // 4)
        one = oracle.Batch([topos[i] for i in pick], [coords[i] for i in pick])
        n2 = np.array([t.natom for t in topos], dtype=np.float64) ** 2
        share = n2[pick].sum() / n2.sum()
        secs1, _ = one.run(repeats=1, nthreads=1)
        single = len(topos) / (secs1 / share)
        cpu = {"value": cpu_value, "unit": "system-evals/s", "cores": threads, "kind": "port",
               "sample": f"all {len(topos)} systems x {reps_cpu} passes ({secs:.1f} s), FP64 "
                         f"energy+force per system, {threads} threads over contiguous system "
                         f"ranges of equal pair work",
               "single_core_value": single,
               "single_core_sample": f"{len(pick)} systems spread evenly over the batch "
                                     f"sorted by N^2 ({100 * share:.1f} % of its pair work), "
                                     f"one pass, {secs1:.1f} s, scaled by pair work",
               "host_cores_visible": visible, "cpu_model": cpu_model(),
               "compiler": "gcc " + " ".join(oracle.CFLAGS)}
        log(f"CPU oracle: {cpu_value:.4e} system-evals/s on {threads} threads "
            f"({single:.4e} on one)")

    extra = {}
    if not args.no_extra:
        import oracle
        extra["lj_table"] = lj_table_workload(S, oracle, topos, coords, max(50, args.steps // 2))
        extra["lj_table"]["combining_rule_kernel_ms"] = kernel_ms
        log(f"LJ tables: {extra['lj_table']}")
        extra["minimization"] = minimization_workload(S, synthetic, oracle, 4096,
                                                      args.min_cycles, not args.no_cpu)
        log(f"minimization: {extra['minimization']}")
        extra["large_batch"] = large_batch_workload(S, synthetic, oracle, 16384,
                                                    max(20, args.steps // 5))
        log(f"large batch: {extra['large_batch']}")
        from stormm_amd import serving
        extra["serving"] = serving.serving_benchmark(seconds=args.serve_seconds, log=log)
        log(f"serving: { {k: v for k, v in extra['serving'].items() if k != 'capacity'} }")
        cap = extra['serving']['capacity']
        log(f"serving capacity: {cap['speedup_vs_batched']:.2f}x the batch minimizer on the same "
            f"conformers, {cap['speedup_vs_static']:.2f}x static serving, "
            f"{cap['speedup_vs_per_request']:.2f}x one batched launch per request")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "system-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"{info.system_count} heterogeneous branched molecules, "
                        "N~U[24,256] atoms, OBC-II GB, one energy+force evaluation of every "
                        "system per step",
            "systems": info.system_count,
            "atoms": atoms,
            "pairs": info.pair_count,
            "valence_terms": info.valence_term_count,
            "force_bits": 23,
            "parallelism": "single GPU, one work unit per block",
        },
        "roofline": {
            "bound": "valu",
            "engine": "FP32 vector ALU: pairwise terms are no dense contraction, so no matrix "
                      "instructions; peak = the FP32 vector datasheet figure",
            "achieved": achieved,
            "peak": roofline.PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / roofline.PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "kernel": "k_mm_eval<float, forces>",
            "kernel_ms": kernel_ms,
            "flop_per_launch": flops,
            "algorithmic_bytes_per_launch": bytes_alg,
            "measured_fp32_fma_tflops": measured_fma,
            "frac_of_measured_fma": achieved / measured_fma if measured_fma > 0 else None,
            "measured_copy_gbps": measured_bw,
        },
        "stage_breakdown": breakdown,
        "fp64_mode": {"kernel_ms": kernel_ms_fp64,
                      "value": info.system_count / (kernel_ms_fp64 * 1.0e-3),
                      "autotuned_shape": shape64},
        "parity": parity,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
