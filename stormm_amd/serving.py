"""Served minimization: a device-resident pool of systems minimized in continuous batches.

A request is one molecule with a set of conformers to minimize (the reference's conformer
application, apps/Conf/src/conformer.cpp:166, minimizes such sets in fixed batches).  The server
admits requests into a MinimizationPool while its device-memory budget allows, advances every
resident conformer a few cycles per tick, and returns a request once all of its conformers have
reached the cycle limit or the RMS-force threshold.  `serve(..., mode="static")` runs the same
requests in plain batches (admit whatever is waiting, minimize the batch to completion, repeat)
for comparison.

    python -m stormm_amd.serving [--seconds 60] [--rate 200] [--budget-mb 256] [--tick 10]

prints one JSON object: the real-time continuous and static runs and the capacity comparison
against the plain batched launches.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import sys
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import PoolConfig, StormmError, check
from .synthesis import _BatchArrays, _shape, precision_code
from .topology import IGB_OBC2, Topology


class MinimizationPool:
    """The C-ABI serving pool: admission against a device-memory budget, resumable cycles."""

    def __init__(self, precision: str = "single", budget_bytes: int = 256 << 20,
                 max_systems: int = 16384, total_cycles: int = 300, sd_cycles: int = 50,
                 dx0: float = 0.01, conv_rms: float = 0.1, igb: int = IGB_OBC2,
                 dielectric: float = 78.5, salt: float = 0.0, coulomb: float = 332.0522172900,
                 shape=None, bits: Optional[Tuple[int, int, int]] = None,
                 clash_cycles: Optional[int] = None, clash_r0: float = 0.2,
                 clash_vdw_ratio: float = 0.5, ntpr: int = 0):
        from . import synthesis as S
        self._L = _lib.load()
        if bits is None:
            bits = (32, 23, 26) if precision == "single" else (32, 40, 32)
        cfg = PoolConfig()
        cfg.precision = precision_code(precision)
        cfg.gpos_bits, cfg.frc_bits, cfg.nrg_bits = bits
        cfg.hbm_budget_bytes = int(budget_bytes)
        cfg.max_systems = int(max_systems)
        cfg.igb = igb
        cfg.dielectric = dielectric
        cfg.salt_concentration = salt
        cfg.coulomb_constant = coulomb
        cfg.minimize = S.minimize_controls(int(total_cycles), int(sd_cycles), float(dx0),
                                           float(conv_rms), clash_cycles, clash_r0,
                                           clash_vdw_ratio, int(ntpr))
        sh = _shape(shape)
        if sh is not None:
            cfg.shape = sh
        h = ctypes.c_void_p()
        check(self._L.stormm_pool_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.config = cfg
        self.nrg_bits = bits[2]

    def __del__(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            self._L.stormm_pool_destroy(self.handle)
            self.handle = None

    def capacity(self) -> Dict[str, int]:
        fa, fs = ctypes.c_int64(0), ctypes.c_int32(0)
        tb, ca = ctypes.c_int64(0), ctypes.c_int64(0)
        check(self._L.stormm_pool_capacity(self.handle, ctypes.byref(fa), ctypes.byref(fs),
                                           ctypes.byref(tb), ctypes.byref(ca)))
        return {"free_atoms": fa.value, "free_slots": fs.value, "template_bytes": tb.value,
                "capacity_atoms": ca.value}

    def admit(self, topos: Sequence[Topology], coords: Sequence[np.ndarray]) -> Optional[List[int]]:
        """Admit systems; returns their slots, or None when the budget cannot hold them."""
        arrs = _BatchArrays(list(topos))
        xyz = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, 3)
                                                   for c in coords]))
        slots = np.zeros(len(topos), dtype=np.int32)
        rc = self._L.stormm_pool_admit(self.handle, ctypes.byref(arrs.struct),
                                       xyz.ctypes.data_as(ctypes.c_void_p),
                                       slots.ctypes.data_as(ctypes.c_void_p))
        if rc == _lib.STORMM_ERR_CAPACITY:
            return None
        check(rc)
        return slots.tolist()

    def admit_conformers(self, topo: Topology,
                         coords: Sequence[np.ndarray]) -> Optional[List[int]]:
        """Admit conformers of one molecule; returns their slots, or None when the budget
        cannot hold them."""
        arrs = _BatchArrays([topo])
        xyz = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.float64).reshape(-1, 3)
                                                   for c in coords]))
        slots = np.zeros(len(coords), dtype=np.int32)
        rc = self._L.stormm_pool_admit_conformers(self.handle, ctypes.byref(arrs.struct),
                                                  len(coords),
                                                  xyz.ctypes.data_as(ctypes.c_void_p),
                                                  slots.ctypes.data_as(ctypes.c_void_p))
        if rc == _lib.STORMM_ERR_CAPACITY:
            return None
        check(rc)
        return slots.tolist()

    def release(self, slots: Sequence[int]) -> None:
        arr = np.ascontiguousarray(np.asarray(slots, dtype=np.int32))
        check(self._L.stormm_pool_release(self.handle, arr.ctypes.data_as(ctypes.c_void_p),
                                          len(arr)))

    def minimize(self, cycles: int, stream=None) -> int:
        """Advance every running system by up to `cycles` cycles; returns how many ran."""
        running = ctypes.c_int32(0)
        check(self._L.stormm_pool_minimize(self.handle, int(cycles), ctypes.byref(running),
                                           None if stream is None else ctypes.c_void_p(stream)))
        return running.value

    def history(self, slot: int):
        """Potential energies (kcal/mol) of one slot's history, one every ntpr cycles from its
        admission, and the ScoreCard rows behind them ([samples][32] int64)."""
        cap = max(1, -(-int(self.config.minimize.total_cycles) //
                       max(1, int(self.config.minimize.diagnostic_print_frequency))))
        nrg = np.zeros(cap)
        rows = np.zeros((cap, 32), dtype=np.int64)
        n = ctypes.c_int32(0)
        vp = ctypes.c_void_p
        check(self._L.stormm_pool_history(self.handle, int(slot), cap, ctypes.byref(n),
                                          nrg.ctypes.data_as(vp), rows.ctypes.data_as(vp)))
        return nrg[:n.value], rows[:n.value]

    def status(self, slots: Sequence[int]):
        """(steps, done, rms force, potential energy) per slot."""
        arr = np.ascontiguousarray(np.asarray(slots, dtype=np.int32))
        n = len(arr)
        steps = np.zeros(n, dtype=np.int32)
        done = np.zeros(n, dtype=np.int32)
        rms = np.zeros(n)
        nrg = np.zeros(n)
        vp = ctypes.c_void_p
        check(self._L.stormm_pool_status(self.handle, arr.ctypes.data_as(vp), n,
                                         steps.ctypes.data_as(vp), done.ctypes.data_as(vp),
                                         rms.ctypes.data_as(vp), nrg.ctypes.data_as(vp)))
        return steps, done.astype(bool), rms, nrg

    def coordinates(self, slot: int, natom: int, fixed: bool = False) -> np.ndarray:
        vp = ctypes.c_void_p
        if fixed:
            out = np.zeros((natom, 3), dtype=np.int64)
            check(self._L.stormm_pool_coordinates(self.handle, int(slot), None,
                                                  out.ctypes.data_as(vp)))
        else:
            out = np.zeros((natom, 3))
            check(self._L.stormm_pool_coordinates(self.handle, int(slot),
                                                  out.ctypes.data_as(vp), None))
        return out


# ----------------------------------------------------------------------------------------------
# Requests
# ----------------------------------------------------------------------------------------------
@dataclass
class Request:
    rid: int
    arrival: float                       # seconds after the start of the run
    topo: Topology
    coords: List[np.ndarray]
    slots: List[int] = field(default_factory=list)
    admitted: float = -1.0
    finished: float = -1.0
    energies: Optional[np.ndarray] = None
    steps: Optional[np.ndarray] = None
    # Per conformer, its potential energies every ntpr cycles (pools with an energy history).
    history: Optional[List[np.ndarray]] = None


def serving_workload(seconds: float, rate: float, seed: int = 3, molecules: int = 256,
                     min_atoms: int = 16, max_atoms: int = 512, max_conformers: int = 64,
                     perturb: float = 0.3) -> List[Request]:
    """Poisson arrivals at `rate` per second for `seconds`; each request holds U{1..64}
    conformers of one molecule with N ~ logU[16, 512] atoms drawn from a fixed set of
    `molecules` seeded molecules, each conformer perturbed by N(0, perturb A) per coordinate."""
    from . import synthetic
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(min_atoms), math.log(max_atoms + 1), molecules))
    sizes = np.clip(sizes.astype(int), min_atoms, max_atoms)
    mol_t, mol_x = synthetic.generate(sizes, [seed * 7919 + i for i in range(molecules)],
                                      gb=True)
    reqs: List[Request] = []
    t = 0.0
    rid = 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > seconds:
            break
        m = int(rng.integers(0, molecules))
        nconf = int(rng.integers(1, max_conformers + 1))
        base = np.asarray(mol_x[m])
        confs = [base + rng.normal(0.0, perturb, base.shape) for _ in range(nconf)]
        reqs.append(Request(rid, t, mol_t[m], confs))
        rid += 1
    return reqs


def serve(pool: MinimizationPool, requests: Sequence[Request], mode: str = "continuous",
          tick_cycles: int = 10, realtime: bool = True, log=None) -> Dict[str, float]:
    """Run the requests through the pool.

    continuous: every tick admits waiting requests while the budget allows, advances every
    resident conformer `tick_cycles` cycles and returns finished requests.
    static: admits whatever is waiting only when the pool is empty, then minimizes that batch
    until all of its conformers are finished.
    With realtime, requests become visible at their arrival times (wall clock); otherwise all are
    waiting from the start."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    waiting: deque = deque()
    active: Dict[int, Request] = {}
    done: List[Request] = []
    nxt = 0
    t0 = time.perf_counter()
    ticks = 0
    busy = 0.0
    while nxt < len(reqs) or waiting or active:
        now = time.perf_counter() - t0
        while nxt < len(reqs) and (not realtime or reqs[nxt].arrival <= now):
            waiting.append(reqs[nxt])
            nxt += 1
        if mode == "continuous" or not active:
            while waiting:
                r = waiting[0]
                slots = pool.admit_conformers(r.topo, r.coords)
                if slots is None:
                    if not active:
                        raise StormmError(_lib.STORMM_ERR_CAPACITY,
                                          f"request {r.rid} does not fit an empty pool")
                    break
                r.slots = slots
                r.admitted = time.perf_counter() - t0
                active[r.rid] = r
                waiting.popleft()
        if not active:
            if nxt < len(reqs):
                time.sleep(max(0.0, reqs[nxt].arrival - (time.perf_counter() - t0)))
            continue
        tb = time.perf_counter()
        pool.minimize(tick_cycles)
        all_slots = [s for r in active.values() for s in r.slots]
        steps, fin, rms, nrg = pool.status(all_slots)
        busy += time.perf_counter() - tb
        ticks += 1
        now = time.perf_counter() - t0
        pos = 0
        finished = []
        for r in active.values():
            k = len(r.slots)
            if fin[pos:pos + k].all():
                r.finished = now
                r.energies = nrg[pos:pos + k].copy()
                r.steps = steps[pos:pos + k].copy()
                if pool.config.minimize.diagnostic_print_frequency > 0:
                    r.history = [pool.history(s)[0] for s in r.slots]
                finished.append(r)
            pos += k
        for r in finished:
            pool.release(r.slots)
            del active[r.rid]
            done.append(r)
        if log is not None and ticks % 50 == 0:
            log(f"[{mode}] t={now:.1f}s ticks={ticks} active={len(active)} "
                f"waiting={len(waiting)} done={len(done)}/{len(reqs)}")
    elapsed = time.perf_counter() - t0
    lat = np.array([r.finished - r.arrival for r in done]) if realtime else \
        np.array([r.finished for r in done])
    conformers = sum(len(r.coords) for r in done)
    cycles = sum(int(r.steps.sum()) for r in done)
    return {"mode": mode, "requests": len(done), "conformers": conformers,
            "elapsed_s": elapsed, "requests_per_s": len(done) / elapsed,
            "minimizations_per_s": conformers / elapsed,
            "cycles_per_conformer": cycles / max(conformers, 1),
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "ticks": ticks, "tick_cycles": tick_cycles, "device_busy_s": busy}


def serve_per_request(requests: Sequence[Request], precision: str = "single",
                      total_cycles: int = 300, sd_cycles: int = 50, dx0: float = 0.01,
                      conv_rms: float = 0.1, realtime: bool = True, persistent: int = 1,
                      shape=None) -> Dict[str, float]:
    """Plain batched launches: each request's conformers are minimized as one batch
    (AtomGraphSynthesis + PhaseSpaceSynthesis + one minimization launch) before the next
    request is taken, the reference's conformer-application pattern."""
    from . import synthesis as S
    bits = (32, 23, 26) if precision == "single" else (32, 40, 32)
    reqs = sorted(requests, key=lambda r: r.arrival)
    mc = S.minimize_controls(total_cycles, sd_cycles, dx0, conv_rms)
    t0 = time.perf_counter()
    lat = []
    conformers = 0
    cycles = 0
    for r in reqs:
        if realtime:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        ag = S.AtomGraphSynthesis([r.topo] * len(r.coords), shape)
        ps = S.PhaseSpaceSynthesis(ag, bits[0], bits[1])
        sc = S.ScoreCard(ag.system_count, bits[2])
        lm = S.LineMinimization(ag.system_count, dx0)
        ps.upload(r.coords)
        S.launch_minimization(precision, ag, ps, mc, sc, lm, persistent=persistent)
        r.energies = sc.potential_energies()
        if persistent:
            r.steps = ag.minimization_status()[0]
        else:
            r.steps = np.full(len(r.coords), total_cycles)
        r.finished = time.perf_counter() - t0
        lat.append(r.finished - (r.arrival if realtime else 0.0))
        conformers += len(r.coords)
        cycles += int(np.sum(r.steps))
    elapsed = time.perf_counter() - t0
    lat = np.array(lat)
    return {"mode": "per_request", "requests": len(reqs), "conformers": conformers,
            "elapsed_s": elapsed, "requests_per_s": len(reqs) / elapsed,
            "minimizations_per_s": conformers / elapsed,
            "cycles_per_conformer": cycles / max(conformers, 1),
            "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None}


def serve_batched(requests: Sequence[Request], precision: str = "single",
                  total_cycles: int = 300, sd_cycles: int = 50, dx0: float = 0.01,
                  conv_rms: float = 0.1, shape=None) -> Dict[str, float]:
    """The batch minimizer on the same conformers: every conformer of every request in one
    AtomGraphSynthesis / PhaseSpaceSynthesis, minimized by one call (the fused cycle kernels
    captured whole in a graph) with the same RMS-force stop and cycle limit.  The synthesis
    build and upload are timed too, as the serving pool's admissions are."""
    from . import synthesis as S
    bits = (32, 23, 26) if precision == "single" else (32, 40, 32)
    mc = S.minimize_controls(total_cycles, sd_cycles, dx0, conv_rms)
    topos = [r.topo for r in requests for _ in r.coords]
    coords = [c for r in requests for c in r.coords]
    S.synchronize()
    t0 = time.perf_counter()
    ag = S.AtomGraphSynthesis(topos, shape)
    ps = S.PhaseSpaceSynthesis(ag, bits[0], bits[1])
    sc = S.ScoreCard(ag.system_count, bits[2])
    lm = S.LineMinimization(ag.system_count, dx0)
    ps.upload(coords)
    S.launch_minimization(precision, ag, ps, mc, sc, lm, persistent=1)
    steps = ag.minimization_status()[0]
    energies = sc.potential_energies()
    elapsed = time.perf_counter() - t0
    pos = 0
    for r in requests:
        k = len(r.coords)
        r.energies = energies[pos:pos + k].copy()
        r.steps = steps[pos:pos + k].copy()
        pos += k
    return {"mode": "batched", "requests": len(requests), "conformers": len(coords),
            "elapsed_s": elapsed, "minimizations_per_s": len(coords) / elapsed,
            "cycles_per_conformer": float(np.mean(steps)) if len(steps) else 0.0}


def serving_benchmark(seconds: float = 5.0, rate: float = 200.0, budget_mb: float = 256.0,
                      tick: int = 10, cycles: int = 300, conv: float = 0.1,
                      baseline_requests: int = 60, log=None) -> Dict[str, object]:
    """The served-minimization workload: Poisson arrivals at `rate` for `seconds`.

    * continuous / static: the request stream in real time (latency under the offered load);
    * capacity: every request of the first `baseline_requests` waiting at the start; continuous
      serving against static serving, against the batch minimizer given all of their conformers
      in one launch (the plain batched launch) and against one batched launch per request."""
    reqs = serving_workload(seconds, rate)

    def fresh(rs):
        return [Request(r.rid, r.arrival, r.topo, r.coords) for r in rs]

    def pool():
        return MinimizationPool(budget_bytes=int(budget_mb * (1 << 20)), total_cycles=cycles,
                                conv_rms=conv)

    out: Dict[str, object] = {"rate_per_s": rate, "arrival_window_s": seconds,
                              "requests": len(reqs),
                              "conformers": sum(len(r.coords) for r in reqs),
                              "hbm_budget_mb": budget_mb, "tick_cycles": tick,
                              "cycles": cycles, "rms_stop": conv}
    # Warm-up: kernels and templates of a few molecules.
    serve(pool(), fresh(reqs[:4]), "continuous", tick, realtime=False)
    out["continuous"] = serve(pool(), fresh(reqs), "continuous", tick, realtime=True, log=log)
    out["static"] = serve(pool(), fresh(reqs), "static", tick, realtime=True, log=log)
    # Capacity with every request waiting at the start: continuous and static serving, the batch
    # minimizer on all of the same conformers at once (the plain batched launch), and one batched
    # launch per request.
    base = reqs[:baseline_requests]
    cap_c = serve(pool(), fresh(base), "continuous", tick, realtime=False)
    cap_s = serve(pool(), fresh(base), "static", tick, realtime=False)
    serve_batched(fresh(base[:2]), total_cycles=cycles, conv_rms=conv)   # warm-up
    cap_b = serve_batched(fresh(base), total_cycles=cycles, conv_rms=conv)
    cap_p = serve_per_request(fresh(base), total_cycles=cycles, conv_rms=conv, realtime=False)
    rate = cap_c["minimizations_per_s"]
    out["capacity"] = {"requests": len(base), "conformers": sum(len(r.coords) for r in base),
                       "continuous_minimizations_per_s": rate,
                       "static_minimizations_per_s": cap_s["minimizations_per_s"],
                       "batched_minimizations_per_s": cap_b["minimizations_per_s"],
                       "per_request_minimizations_per_s": cap_p["minimizations_per_s"],
                       "speedup_vs_batched": rate / cap_b["minimizations_per_s"],
                       "speedup_vs_static": rate / cap_s["minimizations_per_s"],
                       "speedup_vs_per_request": rate / cap_p["minimizations_per_s"],
                       "continuous": cap_c, "static": cap_s, "batched": cap_b,
                       "per_request": cap_p}
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--seconds", type=float, default=60.0, help="arrival window")
    ap.add_argument("--rate", type=float, default=200.0, help="requests per second")
    ap.add_argument("--budget-mb", type=float, default=256.0)
    ap.add_argument("--tick", type=int, default=10, help="cycles per tick")
    ap.add_argument("--cycles", type=int, default=300)
    ap.add_argument("--conv", type=float, default=0.1)
    args = ap.parse_args(argv)
    out = serving_benchmark(args.seconds, args.rate, args.budget_mb, args.tick, args.cycles,
                            args.conv, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
