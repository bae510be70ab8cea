"""Benchmark: stealthy ground states per second on one MI355X.

Headline workload (one "step" = one batch): 256 independent configurations per batch, N log-uniform on [100, 4000],
d in {2, 3} by a fair coin and chi uniform on [0.2, 0.45] (one MT19937 stream seeded 20261203 draws N, d, chi and a
configuration seed per member), each a seeded uniform-random configuration at density 1 with the stealthy constraint
set v(k)=1 for 0 < |k| < K(chi) (GetKs, half space), relaxed with RelaxStructure_NLOPT_Emin (three L-BFGS passes,
history 40, at most 10000 evaluations per pass, Emin = 1e-18) by the device-resident batched solver.  A solve counts
as a ground state when its final energy is below 1e-18.  The timed batches use the same shapes with fresh
configuration seeds; their inputs are generated and admitted (uploaded, resident in HBM) before the timed region.

Prints one JSON line: value = ground states / s over the timed batches, plus E+F evaluations/s, a roofline record
for the E+F contraction kernels (hipEvent-timed on the solver's stream), the other workloads' records (64 seeds of
3D N=2000; 2D N=20000 in FP64 and mixed precision with the error against the oracle; a Poisson request stream with
soft core; the single-solve drop-in path) and a CPU baseline (timed oracle ground states, see cpu_baseline.sample).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_VALU_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak (AMD data sheet)



def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed 256-configuration batches")
    ap.add_argument("--warmup", type=int, default=1, help="untimed batches before them")
    ap.add_argument("--emin", type=float, default=1e-18)
    ap.add_argument("--maxstep", type=int, default=10000)
    ap.add_argument("--history", type=int, default=40)
    ap.add_argument("--check-interval", type=int, default=16)
    ap.add_argument("--slots", type=int, default=0,
                    help="solver slots (0: every timed job at once); fewer slots admit the queued jobs as others finish")
    ap.add_argument("--resident", default="all", choices=["all", "step"],
                    help="all: the timed batches are admitted together and advance in the same launches "
                         "(continuous batching, one tail at the end); step: one batch at a time")
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="bound of the CPU-baseline sample (the stratified ground states stop once it is spent)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sub", action="store_true", help="headline only: skip the other workloads' records")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run the library's copy kernel over this size first (calibrates rocprof FETCH/WRITE_SIZE)")
    ap.add_argument("--graph", action=argparse.BooleanOptionalAction, default=True,
                    help="replay optimizer steps from captured hipGraphs (default; per-launch kernel timing is then "
                         "off, and the kernel roofline comes from a second, stream-launched one-lane solve of the same "
                         "batches)")
    ap.add_argument("--lanes", type=int, default=2, choices=[1, 2],
                    help="solver streams (2, the default: the L-BFGS kernels of one overlap the E+F kernels of the "
                         "other; the kernels then share the GPU, so the roofline comes from the one-lane solve)")
    ap.add_argument("--autotune", action="store_true", help="time candidate tiles per shape class first")
    ap.add_argument("--tiles", type=str, default=None,
                    help="seg2d/seg3d,rho_kernel,rho_chunk,force_chunk,force_ppl for every shape class of the batch "
                         "(0 chunks: automatic), e.g. 32/24,2,0,0,2, instead of the default rule or the autotuner")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="FP64, or FP32 phases and force terms with FP64 sums and an FP64 polish below 1e-8")
    ap.add_argument("--workload", default="batch256",
                    choices=["batch256", "seeds64", "single", "large2d", "mixed2d", "serve", "five", "nufft2d",
                             "targets", "md"],
                    help="batch256: the headline (256 configurations of mixed N, d, chi per batch); seeds64: 64 seeds "
                         "of 3D N=2000 chi=0.45; single: those solves through the single-solve drop-in path; "
                         "large2d: 2D N=20000 chi=0.4; mixed2d: the same in FP64 and in mixed precision, with the "
                         "mixed-precision error; serve: Poisson request stream with soft core; five: multi-species "
                         "ground states of the reference's five-component study (2D N=2000, 5 x 400); nufft2d: "
                         "large 2D systems through the NUFFT backend against the direct sums; targets: target-S(k) "
                         "(equiluminous) relaxes and a variable-cell solve; md: molecular dynamics steps")
    ap.add_argument("--batch", type=int, default=64, help="seeds64 / five / targets: configurations per step")
    ap.add_argument("--requests", type=int, default=2048, help="serve: number of requests")
    ap.add_argument("--serve-slots", type=int, default=1024, help="serve: solver slots (resident requests at most)")
    ap.add_argument("--rate", type=float, default=2000.0,
                    help="serve: arrivals per second (the default saturates the GPU: requests queue)")
    ap.add_argument("--profiler-run", action="store_true",
                    help="for runs under a counter profiler: no hipEvent timing (roofline null), no sub-records")
    ap.add_argument("--profile-evals", type=int, default=0,
                    help="counter-profiler mode: submit one batch, run only this many batched optimizer steps "
                         "(E+F and L-BFGS kernels) and exit; implies --profiler-run and --no-cpu")
    return ap.parse_args()


def copy_probe(mib, repeats=1):
    """The library's plain copy kernel (k_copy_probe) over `mib` MiB: known traffic for the memory counters of a
    profile (mib MiB read and written per launch) and an HBM bandwidth figure.  Returns GB/s (read + write)."""
    import ctypes as C
    from modules_amd._lib import check, lib
    ms = C.c_double(0)
    check(lib().ccgpu_copy_probe(mib << 20, int(repeats), C.byref(ms)))
    gbs = 2 * (mib << 20) / (ms.value * 1e-3) / 1e9
    print(f"copy probe: {mib} MiB read + {mib} MiB written per launch, {ms.value:.3f} ms, {gbs:.0f} GB/s",
          file=sys.stderr, flush=True)
    return gbs


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def _log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def batch_shapes():
    """The 256-configuration batch: (N, d, chi, seed) per configuration, N log-uniform on [100, 4000], d in {2, 3},
    chi uniform on [0.2, 0.45], drawn from one MT19937 stream seeded 20261203; with each shape's constraint set."""
    from modules_amd import GetKs, RandomConfiguration, getK
    from modules_amd.serving import make_batch
    reqs = make_batch()
    ks = [GetKs(RandomConfiguration(0, r.N, r.d), getK(r.chi, 1.0, r.d)) for r in reqs]
    return reqs, ks


def _seed(base, batch_index):
    """Configuration seed of one member of batch `batch_index` (0 = the batch's own draw)."""
    return int((base + 1000003 * batch_index) % (2 ** 31 - 1))


def _roofline(jobs, rho_ms, force_ms, n_rho, n_force):
    """Algorithmic FP64 work of the contraction kernels over the timed launches: per (particle, constraint) pair of
    an E+F evaluation, rho 2d+1 flops and forces 4d+2 (6d+3 in all), summed over every evaluation of every job,
    divided by the summed hipEvent durations of the rho and force launches (on the solver's stream)."""
    r_flops = sum(ev * N * M * (2 * d + 1) for ev, N, M, d in jobs)
    f_flops = sum(ev * N * M * (4 * d + 2) for ev, N, M, d in jobs)
    pairs = sum(ev * N * M for ev, N, M, d in jobs)
    if rho_ms <= 0 or force_ms <= 0:
        return None
    ach_f = f_flops / (force_ms / 1e3) / 1e12
    ach_r = r_flops / (rho_ms / 1e3) / 1e12
    ach_ef = (r_flops + f_flops) / ((rho_ms + force_ms) / 1e3) / 1e12
    P = FP64_VALU_PEAK_TFLOPS
    return {
        "bound": "valu",
        "kernel": "k_force_rows (force contraction, FP64 VALU)",
        "achieved": ach_f, "peak": P, "unit": "TFLOP/s", "frac": ach_f / P,
        "traffic": None,
        "traffic_note": "HBM bytes per launch are measured by rocprofv3 FETCH_SIZE / WRITE_SIZE passes, not in this "
                        "run (profiles/)",
        "avg_launch_ms": force_ms / max(n_force, 1), "launches": n_force,
        "flops_per_pair": "4d+2",
        "rho_kernel": {"kernel": "k_rho_rows", "achieved": ach_r, "frac": ach_r / P,
                       "avg_launch_ms": rho_ms / max(n_rho, 1), "launches": n_rho, "flops_per_pair": "2d+1"},
        "ef_pair": {"achieved": ach_ef, "frac": ach_ef / P, "flops_per_pair": "6d+3"},
        "pairs_per_ef_mean": pairs / max(n_force, 1),
    }


def headline(args, roofline=True):
    """The 256-configuration batch relaxed to E < Emin by the device-resident solver (roofline=False: no second,
    one-lane solve for the kernel roofline)."""
    from modules_amd import RandomConfiguration
    from modules_amd.optimize import BatchSolver
    reqs, ks = batch_shapes()
    B = len(reqs)
    tiles = None
    if args.tiles:
        tiles = record_tiles(reqs, ks, args.tiles)
    elif args.autotune and not args.profiler_run:
        tiles = autotune_classes(reqs, ks)
    K, W = args.steps, args.warmup
    resident_all = args.resident == "all"
    slots = B * (max(K, W, 1) if resident_all else 1)
    if args.slots > 0:
        slots = args.slots
    solver = BatchSolver(max_slots=slots, check_interval=args.check_interval, graph=args.graph,
                         precision=args.precision, lanes=args.lanes)

    def submit(batch_index):
        out = []
        for r, k in zip(reqs, ks):
            cfg = RandomConfiguration(_seed(r.seed, batch_index), r.N, r.d)
            j = solver.submit(cfg, k, np.ones(len(k)), MaxStep=args.maxstep, Emin=args.emin, history=args.history)
            out.append((j, r, len(k)))
        return out

    admit_clock = []

    def solve(groups, clock):
        """Admit the submitted batches (inputs uploaded and resident in HBM), then solve them; only the solve and the
        result read-back are timed (the admission, host-to-HBM uploads included, is timed apart)."""
        ta = time.perf_counter()
        solver.step(0)
        _sync()
        admit_clock.append(time.perf_counter() - ta)
        t0 = time.perf_counter()
        solver.run(chunk=4096)
        res = [[(solver.result(j), r, M) for j, r, M in g] for g in groups]
        _sync()
        clock.append(time.perf_counter() - t0)
        for g in groups:
            for j, _, _ in g:
                solver.release(j)
        return res

    if args.profile_evals > 0:
        # as many batches resident as the timed region holds
        for s in range(K if resident_all else 1):
            submit(s)
        pending = solver.step(args.profile_evals)
        _sync()
        if args.copy_probe_mib > 0:
            copy_probe(args.copy_probe_mib)
        print(json.dumps({"profile_evals": args.profile_evals, "batch": B, "batches": K if resident_all else 1,
                          "pending": pending,
                          "pairs_per_contraction_per_batch": sum(r.N * len(k) for r, k in zip(reqs, ks))}), flush=True)
        return None
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)

    _log(f"warm-up: {W} batch(es) of {B}")
    warm_clock = []
    if resident_all and W > 0:
        solve([submit(1000 + w) for w in range(W)], warm_clock)
    else:
        for w in range(W):
            solve([submit(1000 + w)], warm_clock)
    timing = not args.profiler_run and not args.graph and args.lanes == 1
    solver.set_timing(timing)
    solver.timing(reset=True)
    clock = []
    _log(f"timed: {K} batch(es) of {B}, resident={args.resident}, precision={args.precision}, graph={args.graph}")
    if resident_all:
        groups = [submit(s) for s in range(K)]
        results = [x for g in solve(groups, clock) for x in g]
    else:
        results = []
        for s in range(K):
            results.extend(solve([submit(s)], clock)[0])
    wall = sum(clock)
    admit_s = admit_clock[-1] if admit_clock else 0.0
    rho_ms, force_ms, n_rho, n_force = solver.timing(reset=True)
    launches = solver.launch_stats()
    roof = _roofline([(res["device_evaluations"], r.N, M, r.d) for res, r, M in results], rho_ms, force_ms, n_rho,
                     n_force) if timing else None
    if roof is None and not args.profiler_run and roofline:
        # graph replay or two lanes: per-launch events would miss (graph) or include the other lane's kernels (two
        # lanes), so the kernel roofline comes from a second solve of the same batches on one stream, with stream
        # launches and hipEvents around every contraction launch
        _log("roofline pass: the same batches on one lane with per-launch events")
        solver.set_graph(False)
        solver.set_lanes(1)
        solver.set_timing(True)
        solver.timing(reset=True)
        if resident_all:
            again = [x for g in solve([submit(s) for s in range(K)], []) for x in g]
        else:
            again = [x for s in range(K) for x in solve([submit(s)], [])[0]]
        rho_ms, force_ms, n_rho, n_force = solver.timing(reset=True)
        roof = _roofline([(res["device_evaluations"], r.N, M, r.d) for res, r, M in again], rho_ms, force_ms, n_rho,
                         n_force)
        if roof:
            roof["measured_on"] = ("a second solve of the same batches on one stream (stream launches, hipEvents "
                                   "around each contraction launch); the timed solve ran with graph replay on "
                                   f"{args.lanes} lane(s)")
    solved = sum(1 for res, r, M in results if res["state"] == BatchSolver.DONE and res["energy"] < args.emin)
    evals = np.array([res["device_evaluations"] for res, r, M in results])
    lat = np.array([res["t_done"] - res["t_admit"] for res, r, M in results])
    pairs = sum(r.N * M for res, r, M in results) / max(K, 1)
    return {
        "metric": "stealthy ground states/s at E<1e-18",
        "value": solved / wall,
        "unit": "ground_states/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * wall / max(K, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 phases / f64 sums",
        "data": "synthetic (seeded MT19937 uniform configurations at density 1, GetKs constraint sets, v(k)=1)",
        "config": {"workload": f"batched stealthy relax: {B} configurations per batch, N log-uniform on [100, 4000], "
                               f"d in {{2,3}}, chi uniform on [0.2, 0.45] (master seed 20261203), Emin={args.emin}, "
                               f"3-pass L-BFGS history {args.history}",
                   "batch": B, "pairs_per_contraction_per_batch": pairs,
                   "N_range": [min(r.N for r in reqs), max(r.N for r in reqs)],
                   "resident": args.resident, "graph": bool(args.graph), "precision": args.precision,
                   "lanes": args.lanes, "tiles": tiles or "default rule per system shape"},
        "solved": solved,
        "solves": len(results),
        "launch_bursts": launches,
        "pcie_inclusive_value": solved / (wall + admit_s),
        "admission_s": admit_s,
        "ef_evals_per_s": float(evals.sum()) / wall,
        "device_evals_per_solve_mean": float(evals.mean()),
        "latency_admit_to_done_p50_s": float(np.percentile(lat, 50)),
        "latency_admit_to_done_p99_s": float(np.percentile(lat, 99)),
        "roofline": roof,
    }


def record_tiles(reqs, ks, spec):
    """Record one tile shape (the --tiles string) for every shape class of the batch."""
    from modules_amd import autotune
    segs, rk, rc, fc, ppl = spec.split(",")
    seg2, seg3 = (segs.split("/") + [segs])[:2] if "/" in segs else (segs, segs)
    for r, k in zip(reqs, ks):
        autotune.record(r.N, r.d, len(k), int(seg2 if r.d == 2 else seg3), int(rk), int(rc), int(fc), int(ppl))
    return {"source": "--tiles", "spec": spec}


def autotune_classes(reqs, ks, per_class=1):
    """Time candidate tiles for one configuration of every (d, N, M) shape class of the batch (the engine's tile
    table keys on half-octave buckets of N and M) and record them; returns {class: tiles}."""
    from modules_amd import autotune
    seen = {}
    for r, k in zip(reqs, ks):
        key = (r.d, int(2 * math.log2(r.N)), int(2 * math.log2(max(len(k), 1))))
        if key in seen:
            continue
        seen[key] = autotune.autotune(r.N, r.d, k=k, batch=4, repeats=2)
    return {str(k): v for k, v in seen.items()}


def sub_batch_mixed(args, fp64_value):
    """The headline's batches in mixed precision (FP32 phases and force terms, FP64 sums, FP64 polish below 1e-8 and
    FP64 energies at every stop test), the same timed region and settings, no roofline pass."""
    import copy
    a = copy.copy(args)
    a.precision, a.tiles, a.autotune = "mixed", None, False
    r = headline(a, roofline=False)
    return {"workload": "the headline's batches in mixed precision (FP64 polish below 1e-8)", "value": r["value"],
            "unit": "ground_states/s", "solved": r["solved"], "solves": r["solves"], "ms_per_step": r["ms_per_step"],
            "ef_evals_per_s": r["ef_evals_per_s"], "device_evals_per_solve_mean": r["device_evals_per_solve_mean"],
            "speedup_over_fp64": r["value"] / fp64_value if fp64_value else None}


def sub_seeds64(args, check_oracle=True):
    """64 seeds of 3D N=2000 at chi=0.45 (M=2724), FP64, relaxed side by side; each end point's energy is
    recomputed by the oracle."""
    from modules_amd import GetKs, RandomConfiguration, getK
    from modules_amd.optimize import BatchSolver
    N, d, chi, B = 2000, 3, 0.45, 64
    k = GetKs(RandomConfiguration(0, N, d), getK(chi, 1.0, d))
    v = np.ones(len(k))
    solver = BatchSolver(max_slots=B, check_interval=args.check_interval, lanes=args.lanes)
    cfgs = [RandomConfiguration(s, N, d) for s in range(B)]
    warm = [solver.submit(RandomConfiguration(10 ** 6 + s, N, d), k, v, Emin=args.emin) for s in range(8)]
    res, _ = _solve_all(solver, warm, args.emin)
    for j in warm:
        solver.release(j)
    jobs = [solver.submit(c, k, v, Emin=args.emin) for c in cfgs]
    res, wall = _solve_all(solver, jobs, args.emin)
    solved = sum(1 for r in res if r["energy"] < args.emin)
    out = {"workload": f"{B} seeds x 3D N={N} chi={chi} (M={len(k)}), FP64", "value": solved / wall,
           "unit": "ground_states/s", "solved": solved, "solves": B, "wall_s": wall,
           "device_evals_per_solve_mean": float(np.mean([r["device_evaluations"] for r in res]))}
    if check_oracle:
        from oracle import pyoracle as po
        Eo = [po.total_energy_forces(r["relative"], c.basis, k, v)[0] for r, c in zip(res, cfgs)]
        out["oracle_energy_max"] = float(max(Eo))
        out["oracle_accepted"] = int(sum(1 for e in Eo if e < 10 * args.emin))
    return out


def sub_mixed2d(args):
    """2D N=20000 at chi=0.4: the mixed-precision error at the start configuration against the oracle, then the
    same solve in FP64 and in mixed precision (FP32 phases and force terms, FP64 sums, FP64 polish below 1e-8)."""
    from modules_amd import GetKs, RandomConfiguration, ShiftedCCPotential, getK
    from modules_amd.optimize import BatchSolver
    from oracle import pyoracle as po
    cfg = RandomConfiguration(4, 20000, 2)
    k = GetKs(cfg, getK(0.4, 1.0, 2))
    v = np.ones(len(k))
    Eo, Fo = po.total_energy_forces(cfg.relative, cfg.basis, k, v)
    out = {"workload": f"2D N=20000 chi=0.4 (M={len(k)})", "unit": "ground_states/s"}
    for prec in ("fp64", "mixed"):
        pot = ShiftedCCPotential(2)
        for kk in k:
            pot.AddConstraint(kk, [1.0])
        pot.SetPrecision(prec)
        pot.SetConfiguration(cfg)
        E, F = pot.EnergyAndAllForce()
        solver = BatchSolver(max_slots=1, check_interval=args.check_interval, precision=prec)
        job = solver.submit(cfg, k, v, MaxStep=args.maxstep, Emin=args.emin)
        res, wall = _solve_all(solver, [job], args.emin)
        r = res[0]
        out[prec] = {"start_vs_oracle": {"dE_over_E": abs(E - Eo) / abs(Eo),
                                         "max_dF_over_max_F": float(np.max(np.abs(F - Fo)) / np.max(np.abs(Fo)))},
                     "value": (1.0 if r["energy"] < args.emin else 0.0) / wall, "wall_s": wall,
                     "energy": r["energy"], "objective_evals": r["evaluations"],
                     "device_evals": r["device_evaluations"], "switch_eval": r["switch_eval"],
                     "ms_per_eval": 1e3 * wall / max(r["device_evaluations"], 1)}
        out[prec]["oracle_energy"] = po.total_energy_forces(r["relative"], cfg.basis, k, v)[0]
    out["speedup_mixed_over_fp64"] = out["fp64"]["wall_s"] / out["mixed"]["wall_s"]
    return out


def sub_serving(args, requests=None, rate=None):
    """Poisson request stream (N lognormal median 1000, sigma_ln 0.6, clipped to [100, 50000]; d in {2,3}; chi
    uniform on [0.1, 0.49]; soft core e0 = 1, sigma = 0.4 density^(-1/d)), arrivals faster than one at a time can be
    served, on one continuous-batching solver; latency = completion - arrival."""
    from modules_amd.serving import make_stream, serve
    requests = requests or args.requests
    rate = rate or args.rate
    serve(make_stream(8, rate, seed=99), Emin=args.emin, MaxStep=args.maxstep, max_slots=64)  # warm
    reqs, wall = serve(make_stream(requests, rate), Emin=args.emin, MaxStep=args.maxstep, max_slots=args.serve_slots,
                       graph=args.graph)
    lat = np.array([r.done - r.arrival for r in reqs])
    solved = sum(1 for r in reqs if r.result.get("energy", 1.0) < args.emin)
    return {"workload": f"serving: {requests} Poisson requests at {rate}/s, soft core e0=1 sigma=0.4",
            "value": solved / wall, "unit": "ground_states/s", "solved": solved, "requests": len(reqs),
            "wall_s": wall, "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)), "N_median": float(np.median([r.N for r in reqs])),
            "N_max": int(max(r.N for r in reqs)), "last_arrival_s": float(max(r.arrival for r in reqs))}


def sub_single(args, seeds=2):
    """The single-solve drop-in path (host L-BFGS in NLopt order, one GPU E+F per objective evaluation through the
    Potential interface) on seeds of 3D N=2000 chi=0.45."""
    from modules_amd import GetKs, RandomConfiguration, ShiftedCCPotential, getK
    from modules_amd.optimize import RelaxStructure_NLOPT_Emin
    N, d, chi = 2000, 3, 0.45
    cfgs = [RandomConfiguration(s, N, d) for s in range(seeds)]
    k = GetKs(cfgs[0], getK(chi, 1.0, d))
    pot = ShiftedCCPotential(d)
    for kk in k:
        pot.AddConstraint(kk, [1.0])
    RelaxStructure_NLOPT_Emin(RandomConfiguration(999999, N, d), pot, MaxStep=200, Emin=args.emin)  # warm
    t0 = time.perf_counter()
    res = [RelaxStructure_NLOPT_Emin(c, pot, MaxStep=args.maxstep, Emin=args.emin) for c in cfgs]
    wall = time.perf_counter() - t0
    solved = sum(1 for E, ev, st in res if E < args.emin)
    evals = sum(ev for E, ev, st in res)
    return {"workload": f"single-solve drop-in path: {seeds} seeds x 3D N={N} chi={chi} (M={len(k)})",
            "value": solved / wall, "unit": "ground_states/s", "solved": solved, "solves": seeds, "wall_s": wall,
            "ms_per_objective_eval": 1e3 * wall / max(evals, 1), "latency_mean_s": wall / seeds}


def _cpu_topology():
    """Threads this process may run on, and the physical cores under them (sysfs topology)."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        try:
            base = f"/sys/devices/system/cpu/cpu{c}/topology/"
            cores.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    return len(cpus), len(cores)


def cpu_baseline_batch(reqs, ks, args, strata=8):
    """Ground states timed on the host CPU: the oracle (the reference direct path restated in C, OpenMP over
    constraints and particles, libm sin/cos per pair) driven by the reference's own LuksanPLIS (compiled from
    potential/LuksanPLIS.cpp into oracle/_ref) through the three-pass RelaxStructure_NLOPT_Emin.  The sample is
    stratified over the batch: the 256 configurations are sorted by N*M and cut into `strata` equal groups, and the
    middle configuration of each group is solved, cheapest group first, until --cpu-seconds is spent.  With equal
    groups the batch's CPU rate is (solved in the sample) / (sample seconds) over the groups covered; the record
    says how many."""
    os.environ.setdefault("OMP_WAIT_POLICY", "ACTIVE")
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from modules_amd import GetKs, RandomConfiguration, getK
    from oracle import pyoracle as po
    threads_avail, phys = _cpu_topology()
    threads = min(16, threads_avail)
    order = sorted(range(len(reqs)), key=lambda i: reqs[i].N * len(ks[i]))
    g = len(order) // strata
    picks = [order[q * g + g // 2] for q in range(strata)]
    sample, spent = [], 0.0
    for i in picks:
        if spent >= args.cpu_seconds:
            break
        r, k = reqs[i], ks[i]
        cfg = RandomConfiguration(r.seed, r.N, r.d)
        t0 = time.perf_counter()
        E, s, ev, st = po.relax_emin(cfg.relative, cfg.basis, k, np.ones(len(k)), maxstep=args.maxstep,
                                     Emin=args.emin, nthreads=threads)
        dt = time.perf_counter() - t0
        spent += dt
        sample.append({"N": r.N, "d": r.d, "chi": round(r.chi, 4), "M": len(k), "seconds": dt, "evaluations": ev,
                       "energy": E, "solved": bool(E < args.emin)})
    solved = sum(1 for x in sample if x["solved"])
    secs = sum(x["seconds"] for x in sample)
    # anchor: the reference's published 2D N=500 chi=0.2996 E+F on 2 threads (4.0535 ms,
    # docs/finufft/reports/one-component-benchmark.md:37)
    acfg = RandomConfiguration(0, 500, 2)
    ak = GetKs(acfg, getK(0.2996, 1.0, 2))
    ta = _time_ef(po, acfg, ak, np.ones(len(ak)), 2, 1.0, min_n=30)
    covered = len(sample)
    return {"value": solved / secs if secs > 0 else None, "unit": "ground_states/s", "cores": threads, "kind": "port",
            "sample": f"{covered} of {strata} strata of the 256-configuration batch (sorted by N*M, middle member of "
                      f"each stratum, cheapest stratum first): timed oracle ground states (RelaxStructure_NLOPT_Emin "
                      f"with the reference LuksanPLIS, oracle E+F) on {threads} threads, {secs:.1f} s",
            "strata_covered": covered, "strata": strata,
            "strata_note": "when fewer strata than all are covered the sample leans to cheap configurations, so the "
                           "rate overstates the CPU",
            "solves": sample,
            "threads": threads, "physical_cores": phys,
            "anchor_2d_N500_chi0.2996_2threads": {"median_ms_per_ef": 1e3 * float(np.median(ta)), "samples": len(ta),
                                                  "constraints": len(ak), "published_ms": 4.0535},
            "cpu_model": _cpu_model(),
            "omp": {"OMP_WAIT_POLICY": os.environ.get("OMP_WAIT_POLICY"),
                    "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND")}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); the path shards by configuration "
              "(each GPU relaxes its own configurations, with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.profile_evals > 0:
        args.profiler_run = True
        args.no_cpu = True
    if args.workload != "batch256":
        print(json.dumps(OTHER_WORKLOADS[args.workload](args)), flush=True)
        return
    out = headline(args)
    if out is None:
        return
    if not args.profiler_run and not args.no_sub:
        subs = {}
        for name, fn in (("batch256_mixed", lambda a: sub_batch_mixed(a, out["value"])),
                         ("seeds64_3d_N2000", sub_seeds64), ("mixed_2d_N20000", sub_mixed2d),
                         ("serving", sub_serving),
                         # below the saturating rate: the latency of a request rather than of a queue
                         ("serving_200_per_s", lambda a: sub_serving(a, requests=400, rate=200.0)),
                         ("single_solve_3d_N2000", sub_single)):
            _log(f"sub-record {name}")
            subs[name] = fn(args)
        out["workloads"] = subs
        if subs["single_solve_3d_N2000"]["value"] and subs["seeds64_3d_N2000"]["value"]:
            out["batched_over_single_solve_3d_N2000"] = (subs["seeds64_3d_N2000"]["value"] /
                                                         subs["single_solve_3d_N2000"]["value"])
    if not args.no_cpu:
        _log("CPU baseline")
        reqs, ks = batch_shapes()
        out["cpu_baseline"] = cpu_baseline_batch(reqs, ks, args)
        cb = out["cpu_baseline"]["value"]
        out["gpu_over_cpu"] = out["value"] / cb if cb else None
    print(json.dumps(out), flush=True)


def _tune(args, cfg, k):
    """Autotuned tiles for one configuration of this shape (with --autotune)."""
    if not args.autotune:
        return None
    from modules_amd import autotune
    return autotune.autotune(cfg.NumParticle(), cfg.GetDimension(), k=k, basis=cfg.basis, batch=1, repeats=2)


def _solve_all(solver, jobs, emin):
    """Admit the submitted jobs (inputs resident in HBM), then run the solver until they finish; returns results
    and the wall seconds of the solve and readback."""
    solver.step(0)
    t0 = time.perf_counter()
    solver.run(chunk=1024)
    wall = time.perf_counter() - t0
    res = [solver.result(j) for j in jobs]
    return res, wall


def workload_large2d(args):
    """2D N=20000 at density 1, chi=0.4 (M ~ 16000): one configuration relaxed to E<1e-18 in FP64."""
    from modules_amd import GetKs, RandomConfiguration, getK
    from modules_amd.optimize import BatchSolver
    cfg = RandomConfiguration(4, 20000, 2)
    k = GetKs(cfg, getK(0.4, 1.0, 2))
    tiles = _tune(args, cfg, k)
    solver = BatchSolver(max_slots=1, check_interval=args.check_interval)
    job = solver.submit(cfg, k, np.ones(len(k)), MaxStep=args.maxstep, Emin=args.emin)
    res, wall = _solve_all(solver, [job], args.emin)
    r = res[0]
    return {"metric": "stealthy ground states/s at E<1e-18", "value": (1.0 if r["energy"] < args.emin else 0.0) / wall,
            "unit": "ground_states/s", "workload": f"2D N=20000 chi=0.4 (M={len(k)}), FP64", "tiles": tiles,
            "energy": r["energy"],
            "objective_evals": r["evaluations"], "device_evals": r["device_evaluations"], "wall_s": wall,
            "ef_evals_per_s": r["device_evaluations"] / wall, "dtype": "f64"}


def _shell_kset(cfg, target):
    """Wavevectors of the complete |k| shells whose count is nearest to `target` (GetKs over a K between two
    shells)."""
    from modules_amd import GetKs
    k_all = GetKs(cfg, 2.0 * math.pi * math.sqrt(2.0 * target / math.pi) / cfg.basis[0, 0] * 4.0)
    m2 = np.round(np.sum(k_all ** 2, axis=1), 9)
    shells = np.unique(m2)
    counts = np.array([np.sum(m2 <= q) for q in shells])
    i = int(np.argmin(np.abs(counts - target)))
    K = math.sqrt(0.5 * (shells[i] + shells[i + 1]))
    k = GetKs(cfg, K)
    assert len(k) == counts[i]
    return k


def five_component_setup(seed, N=2000, species=5, chi=0.10, phi=0.2):
    """The reference's five-component study (docs/finufft/reports/five-component-original.md): 2D, N=2000 (400
    per species, contiguous blocks as GetMultiCCO's random mode assigns them, apps/CCOptimization/main.cpp:
    1195-1212), equal radii sqrt(phi/pi) (total packing fraction 0.2 at density 1), u0 = e0 = 1, every species
    constrained on the complete shells nearest chi_i = M / (2 (N_s - 1))."""
    from modules_amd import MultiStealthyPotential, RandomConfiguration
    cfg = RandomConfiguration(seed, N, 2)
    n_s = N // species
    k = _shell_kset(cfg, chi * 2 * (n_s - 1))
    pot = MultiStealthyPotential(2, 1.0)
    pot.SetConfiguration(cfg)
    pot.SetNumSpecies(species)
    pot.SetSecureRadii(np.full(N, math.sqrt(phi / math.pi)))
    for q in range(species):
        for i in range(q * n_s, (q + 1) * n_s):
            pot.SetSpeciesOfParticle(i, q)
        for kk in k:
            pot.AddConstraint(q, kk, [1.0])
    return cfg, pot, k


def workload_five(args):
    """Multi-species ground states (MultiStealthyPotential): the reference's five-component study shape, --batch
    random starts relaxed side by side with RelaxStructure_NLOPT (three passes, at most 2000 evaluations each, the
    study's limits) and accepted at E <= 1e-11 (its acceptance); plus the single-solve drop-in path on a few
    starts.  The study's direct-sum solves took 6.46, 6.61 and 5.77 s per configuration on 4 threads
    (docs/finufft/reports/five-component-original.md:17-21).  (The oracle is no CPU baseline here: its contact
    term is an all-pairs image sum, not the reference's cell-list walk.)"""
    from modules_amd.optimize import BatchSolver
    B = args.batch
    maxstep = 2000
    setups = [five_component_setup(s) for s in range(B)]
    M = len(setups[0][2])
    solver = BatchSolver(max_slots=max(1, B), check_interval=args.check_interval, lanes=args.lanes)
    jobs = [solver.submit_multi(cfg, pot, MaxStep=maxstep) for cfg, pot, _ in setups]
    res, wall = _solve_all(solver, jobs, 1e-11)
    energies = np.array([r["energy"] for r in res])
    accepted = int(np.sum(energies <= 1e-11))
    lat = np.array([r["t_done"] - r["t_admit"] for r in res])
    # single-solve drop-in path (host L-BFGS, one GPU E+F per objective evaluation) on the first starts
    n_single = min(3, B)
    single = []
    for cfg, pot, _ in setups[:n_single]:
        pot.SetConfiguration(cfg)
        t0 = time.perf_counter()
        E, ev, st = pot.Relax(maxstep=maxstep)
        single.append((time.perf_counter() - t0, E, ev))
    evals = np.array([r["device_evaluations"] for r in res])
    return {"metric": "multi-species ground states/s at E<=1e-11", "value": accepted / wall,
            "unit": "ground_states/s",
            "workload": f"five-component 2D N=2000 (5 x 400), chi_i={M}/798, phi=0.2, u0=e0=1: {B} starts "
                        "relaxed side by side",
            "accepted": accepted, "solves": B, "wall_s": wall, "energy_max": float(energies.max()),
            "latency_mean_s": float(lat.mean()), "latency_p99_s": float(np.percentile(lat, 99)),
            "device_evals_mean": float(evals.mean()), "ef_evals_per_s": float(evals.sum() / wall),
            "single_solve": {"seconds": [x[0] for x in single], "energy": [x[1] for x in single],
                             "evaluations": [x[2] for x in single]},
            "published_cpu_s_per_config_4_threads": [6.459, 6.608, 5.772],
            "vs_published_4_thread_rate": (accepted / wall) / (1.0 / 6.28), "dtype": "f64"}


def workload_nufft2d(args):
    """Large 2D systems through the NUFFT backend (Type 1 rho, Type 2 forces) against the direct sums: 2D N=20000
    chi=0.4 solved both ways (with the NUFFT error at the start configuration against the direct path, which the
    tests pin to the oracle), then 2D N=100000 chi=0.4 (M ~ 80000) solved through the NUFFT, with the direct path's
    E+F time per evaluation beside it."""
    from modules_amd import GetKs, RandomConfiguration, ShiftedCCPotential, getK
    from modules_amd.optimize import BatchSolver

    def solve(cfg, k, backend):
        solver = BatchSolver(max_slots=1, check_interval=args.check_interval, rho_backend=backend)
        job = solver.submit(cfg, k, np.ones(len(k)), MaxStep=args.maxstep, Emin=args.emin)
        res, wall = _solve_all(solver, [job], args.emin)
        r = res[0]
        return {"wall_s": wall, "energy": r["energy"], "objective_evals": r["evaluations"],
                "device_evals": r["device_evaluations"], "ms_per_eval": 1000 * wall / max(r["device_evaluations"], 1)}

    def ef_ms(cfg, k, backend, reps):
        pot = ShiftedCCPotential(2)
        for kk in k:
            pot.AddConstraint(kk, [1.0])
        pot.SetRhoBackend(backend)
        pot.SetConfiguration(cfg)
        E, F = pot.EnergyAndAllForce()
        t0 = time.perf_counter()
        for _ in range(reps):
            pot.SetConfiguration(cfg)
            pot.EnergyAndAllForce()
        return 1000 * (time.perf_counter() - t0) / reps, E, F, pot.GetLastRhoBackendUsed()

    def gate(cfg, k):
        # what the gated "nufft" backend picks for this shape, and the cost model's predictions behind it
        pot = ShiftedCCPotential(2)
        for kk in k:
            pot.AddConstraint(kk, [1.0])
        pot.SetRhoBackend("nufft")
        pot.SetConfiguration(cfg)
        why, pd, pn = pot.NufftGateReason()
        return {"backend": pot.GetLastRhoBackendUsed(), "reason": why, "predicted_us_direct": pd,
                "predicted_us_nufft": pn}

    def device_ms(cfg, k):
        # one E+F of each path on the device (hipEvents, mean of 5; the host copies of the potential path excluded);
        # the timed choice it records is forgotten again so that the gate reported above stays the cost model's
        from modules_amd import autotune
        use, md, mn = autotune.autotune_backend(cfg.NumParticle(), 2, k=np.asarray(k), basis=cfg.basis, repeats=5)
        autotune.clear()
        return md, mn

    out = {"metric": "stealthy ground states/s at E<1e-18", "unit": "ground_states/s", "dtype": "f64"}
    cfg = RandomConfiguration(4, 20000, 2)
    k = GetKs(cfg, getK(0.4, 1.0, 2))
    tiles = _tune(args, cfg, k)
    g = gate(cfg, k)
    ms_d, E_d, F_d, _ = ef_ms(cfg, k, "direct", 20)
    ms_n, E_n, F_n, _ = ef_ms(cfg, k, "nufft_always", 20)
    dev_d, dev_n = device_ms(cfg, k)
    out["N20000"] = {"M": len(k), "tiles": tiles, "gate": g,
                     "ef_ms_direct_device": dev_d, "ef_ms_nufft_device": dev_n,
                     "ef_ms_direct_host_path": ms_d, "ef_ms_nufft_host_path": ms_n,
                     "nufft_vs_direct_start": {"dE_over_E": abs(E_n - E_d) / abs(E_d),
                                               "max_dF_over_max_F": float(np.max(np.abs(F_n - F_d)) / np.max(np.abs(F_d)))},
                     "direct": solve(cfg, k, "direct"), "nufft": solve(cfg, k, "nufft_always")}
    cfg = RandomConfiguration(5, 100000, 2)
    k = GetKs(cfg, getK(0.4, 1.0, 2))
    g = gate(cfg, k)
    ms_d, E_d, F_d, _ = ef_ms(cfg, k, "direct", 2)
    ms_n, E_n, F_n, _ = ef_ms(cfg, k, "nufft_always", 10)
    dev_d, dev_n = device_ms(cfg, k)
    r = solve(cfg, k, "nufft")
    out["N100000"] = {"M": len(k), "gate": g, "ef_ms_direct_device": dev_d, "ef_ms_nufft_device": dev_n,
                      "ef_ms_direct_host_path": ms_d, "ef_ms_nufft_host_path": ms_n,
                      "nufft_vs_direct_start": {"dE_over_E": abs(E_n - E_d) / abs(E_d),
                                                "max_dF_over_max_F": float(np.max(np.abs(F_n - F_d)) / np.max(np.abs(F_d)))},
                      "nufft": r,
                      "direct_solve_estimate_s": (dev_d / 1000 + r["wall_s"] / max(r["device_evals"], 1)
                                                  - dev_n / 1000) * r["device_evals"]}
    out["value"] = (1.0 if r["energy"] < args.emin else 0.0) / r["wall_s"]
    out["workload"] = "2D N=100000 chi=0.4 through the NUFFT (and N=20000 both ways)"
    return out


def workload_targets(args):
    """Target-S(k) and variable-cell rows.  (1) Equiluminous patterns (RepulsiveCCPotential with S0 != 0,
    apps/CCOptimization/main.cpp:302-318): 2D N=2000, every constraint with 0 < |k| <= K(chi=0.3) weighted 1 with
    the target S(k) = 0.5, --batch random starts relaxed side by side (three passes, at most 2000 evaluations each);
    the oracle's S2 E+F (rho and forces restated in C) timed on the host threads beside it.  (2) A variable-cell
    stealthy solve (K2Potential, CCPotential_Kernels.h:191): 3D N=2000 with rK = 1 at the density giving M ~ 0.3 d N
    constraints, RelaxStructure_NLOPT with Switch 2 (cell and particles) at P = 1e-4, each evaluation regenerating
    the constraints for the new cell; milliseconds per evaluation."""
    from modules_amd import Configuration, GetKs, RandomConfiguration, getK
    from modules_amd import varbox as VB
    from modules_amd.optimize import BatchSolver, RelaxStructure_NLOPT
    from oracle import pyoracle as po
    out = {"metric": "equiluminous (target-S(k)) relaxes/s", "unit": "relaxes/s", "dtype": "f64"}
    B, N, d = args.batch, 2000, 2
    cfgs = [RandomConfiguration(100 + i, N, d) for i in range(B)]
    k = GetKs(cfgs[0], getK(0.3, 1.0, d))
    v = np.ones(len(k))
    s0 = np.full(len(k), 0.5)
    solver = BatchSolver(max_slots=max(1, B), check_interval=args.check_interval, lanes=args.lanes)
    jobs = [solver.submit_s2(c, k, v, s0, MaxStep=2000) for c in cfgs]
    res, wall = _solve_all(solver, jobs, -1e10)
    E = np.array([r["energy"] for r in res])
    evals = np.array([r["device_evaluations"] for r in res])
    out["value"] = B / wall
    out["s2"] = {"M": len(k), "solves": B, "wall_s": wall, "energy_max": float(E.max()),
                 "energy_median": float(np.median(E)), "device_evals_mean": float(evals.mean()),
                 "ef_evals_per_s": float(evals.sum() / wall)}
    # oracle S2 E+F on the host threads, bounded sample
    threads = min(16, os.cpu_count() or 1)
    c0 = cfgs[0]
    t = []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 8.0 and len(t) < 500:
        t1 = time.perf_counter()
        re, im = po.rho(c0.relative, c0.basis, k, nthreads=threads)
        po.s2_energy(v, s0, re, im, N)
        po.s2_forces(c0.relative, c0.basis, k, v, s0, re, im, nthreads=threads)
        t.append(time.perf_counter() - t1)
    ef = len(t) / sum(t)
    out["cpu_baseline"] = {"value": ef / max(float(evals.mean()), 1.0), "unit": "relaxes/s", "cores": threads,
                           "kind": "port", "sample": f"{len(t)} oracle S2 E+F evaluations of one N={N} d={d} "
                           f"configuration on {threads} threads ({ef:.3g} E+F/s) divided by "
                           f"{evals.mean():.0f} evaluations per relax"}
    # variable-cell stealthy solve, Switch 2
    d3, N3 = 3, 2000
    L = (0.3 * d3 * N3 * 12 * math.pi ** 2) ** (1.0 / 3)  # M ~ V / (12 pi^2) points with |k| < 1 (one of +-k)
    c3 = Configuration(np.diag([L, L, L]), RandomConfiguration(7, N3, d3).relative)
    pot = VB.K2Potential(d3, 1.0)
    pot.SetConfiguration(c3)
    M3 = len(pot.constraints)
    H0 = pot.Energy() + 1e-4 * c3.PeriodicVolume()
    t0 = time.perf_counter()
    H, ev, st = RelaxStructure_NLOPT(c3, pot, 1e-4, 2, 0.0, 400)
    wall3 = time.perf_counter() - t0
    out["varbox"] = {"workload": f"K2Potential 3D N={N3}, rK=1, start cell {L:.2f}^3 (M={M3}), P=1e-4, Switch 2",
                     "H0": H0, "H": H, "evaluations": ev, "status": st, "wall_s": wall3,
                     "ms_per_eval": 1000 * wall3 / max(ev, 1), "final_M": len(pot.constraints),
                     "final_volume": c3.PeriodicVolume()}
    out["workload"] = (f"equiluminous S(k)=0.5 relaxes, 2D N={N} chi=0.3 (M={len(k)}), {B} starts; variable-cell "
                       f"3D N={N3} K2Potential Switch 2")
    return out


def workload_md(args):
    """Molecular dynamics (ParticleMolecularDynamics, the MD mode of apps/CCOptimization/main.cpp:1875-2010) on
    the headline system, 3D N=2000 chi=0.45 (M=2724) with the soft core e0=1, sigma=0.5: velocity Verlet steps
    resident on the GPU (one E+F each), Anderson steps, and the oracle's E+F on the host threads beside it (one
    E+F per MD step)."""
    from modules_amd import GetKs, RandomConfiguration, RepulsiveCCPotential, getK
    from modules_amd.md import ParticleMolecularDynamics
    from modules_amd.potential import RandomGenerator
    from oracle import pyoracle as po
    cfg = RandomConfiguration(0, 2000, 3)
    k = GetKs(cfg, getK(0.45, 1.0, 3))
    pot = RepulsiveCCPotential(3, 1.0, 0.5)
    for kk in k:
        pot.CCPotential.AddConstraint(kk, [1.0])
    pot.SetConfiguration(cfg)
    md = ParticleMolecularDynamics(cfg, 0.001, 1.0)
    gen = RandomGenerator(1)
    md.SetRandomSpeed(0.01, gen)
    md.Evolve(20, pot)  # warm
    steps = 2000
    t0 = time.perf_counter()
    md.Evolve(steps, pot)
    nve = time.perf_counter() - t0
    t0 = time.perf_counter()
    md.AndersonEvolve(steps, pot, 0.01, 0.01, gen)
    anderson = time.perf_counter() - t0
    threads = min(16, os.cpu_count() or 1)
    t = _time_ef(po, cfg, k, np.ones(len(k)), threads, 8.0)
    ef = len(t) / t.sum()
    return {"metric": "MD steps/s", "value": steps / nve, "unit": "steps/s", "dtype": "f64",
            "workload": f"velocity Verlet MD, 3D N=2000 chi=0.45 (M={len(k)}), soft core e0=1 sigma=0.5, dt=0.001",
            "nve": {"steps": steps, "wall_s": nve, "ms_per_step": 1000 * nve / steps},
            "anderson": {"steps": steps, "wall_s": anderson, "ms_per_step": 1000 * anderson / steps},
            "cpu_baseline": {"value": ef, "unit": "steps/s", "cores": threads, "kind": "port",
                             "sample": f"{len(t)} oracle CC E+F evaluations (soft core left out) of the start configuration "
                                       f"on {threads} threads: one per MD step"}}


OTHER_WORKLOADS = {"seeds64": sub_seeds64, "single": sub_single, "large2d": workload_large2d,
                   "mixed2d": sub_mixed2d, "serve": sub_serving, "five": workload_five, "nufft2d": workload_nufft2d,
                   "targets": workload_targets, "md": workload_md}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_ef(po, cfg, k, v, threads, seconds, max_n=2000, min_n=1):
    """Warm E+F evaluations of one configuration on `threads` threads for about `seconds`; per-call times."""
    po.total_energy_forces(cfg.relative, cfg.basis, k, v, nthreads=threads)  # warm
    times = []
    t0 = time.perf_counter()
    while True:
        t = time.perf_counter()
        po.total_energy_forces(cfg.relative, cfg.basis, k, v, nthreads=threads)
        times.append(time.perf_counter() - t)
        if (time.perf_counter() - t0 >= seconds and len(times) >= min_n) or len(times) >= max_n:
            break
    return np.array(times)


if __name__ == "__main__":
    main()
