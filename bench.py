"""Benchmark: stealthy ground states per second on one MI355X.

Workload (one "step"): a batch of 64 seeded configurations of N=2000 uniform-random points in a periodic cube
at density 1 (3D), stealthy constraint set v(k)=1 for 0 < |k| < K with K = getK(chi=0.45, 1, 3) (GetKs, half
space), each relaxed with RelaxStructure_NLOPT_Emin (three L-BFGS passes, history 40, at most 10000 evaluations
per pass, Emin = 1e-18) by the device-resident batched solver.  A solve counts as a ground state when its final
energy is below 1e-18.  Inputs are synthetic (seeded MT19937), generated before the timed region.

Prints one JSON line: value = ground states / s over the timed steps, plus E+F evaluations/s, p50/p99 solve
latency, a roofline record for the E+F contraction kernels (hipEvent-timed on the solver's stream) and a CPU
baseline (the oracle's OpenMP E+F restatement timed on this host, see cpu_baseline.sample).
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

# HBM bytes per launch of the force kernel at the headline shape (64 x 3D N=2000, M=2724, tiles 24,0,128,32,1),
# from rocprofv3 FETCH_SIZE and WRITE_SIZE passes scaled by the library copy kernel's counter factors; the
# passes, the commands and the factors are in profiles/counters_ef_lbfgs_batched64_3d_N2000_chi045.md
FORCE_TRAFFIC_BYTES = {(3, 2000, 2724, 64): {"read": 43.5e6, "write": 24.2e6}}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64, help="configurations per step")
    ap.add_argument("--N", type=int, default=2000)
    ap.add_argument("--dim", type=int, default=3)
    ap.add_argument("--chi", type=float, default=0.45)
    ap.add_argument("--emin", type=float, default=1e-18)
    ap.add_argument("--maxstep", type=int, default=10000)
    ap.add_argument("--history", type=int, default=40)
    ap.add_argument("--check-interval", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="bound of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one torch device copy of this size first (calibrates rocprof FETCH/WRITE_SIZE)")
    ap.add_argument("--graph", action="store_true", help="replay optimizer steps from a captured hipGraph")
    ap.add_argument("--lanes", type=int, default=1, choices=[1, 2],
                    help="solver streams (2: the L-BFGS kernels of one overlap the E+F kernels of the other; the "
                         "kernels then share the GPU and their event durations no longer give a per-kernel roofline)")
    ap.add_argument("--no-autotune", action="store_true", help="skip the tile autotuner (default tiles)")
    ap.add_argument("--tiles", type=str, default=None,
                    help="seg,rho_kernel,rho_chunk,force_chunk,force_ppl: use these tiles instead of autotuning")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="batched / mixed256: FP64, or FP32 phases and force terms with FP64 sums and an FP64 "
                         "polish below 1e-8")
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "single", "mixed256", "large2d", "mixed2d", "serve", "five", "nufft2d",
                             "targets", "md"],
                    help="batched: the headline (64 seeds, 3D N=2000 chi=0.45); single: the same solves through the "
                         "single-solve drop-in path; mixed256: 256 configurations of mixed N, d, chi; large2d: 2D "
                         "N=20000 chi=0.4; mixed2d: the same solved in FP64 and in mixed precision, with the "
                         "mixed-precision error; serve: Poisson request stream with soft core; five: multi-species "
                         "ground states of the reference's five-component study (2D N=2000, 5 x 400); nufft2d: "
                         "large 2D systems through the NUFFT backend against the direct sums; targets: target-S(k) "
                         "(equiluminous) relaxes and a variable-cell solve; md: molecular dynamics steps")
    ap.add_argument("--requests", type=int, default=48, help="serve: number of requests")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: arrivals per second")
    ap.add_argument("--profiler-run", action="store_true",
                    help="for runs under a counter profiler: no autotuner and no hipEvent timing (roofline null)")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); the path shards by configuration "
              "(each GPU relaxes its own configurations, with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)

    from modules_amd import GetKs, RandomConfiguration, getK
    from modules_amd.optimize import BatchSolver

    if args.workload != "batched":
        print(json.dumps(OTHER_WORKLOADS[args.workload](args)), flush=True)
        return

    N, d, chi = args.N, args.dim, args.chi
    K = getK(chi, 1.0, d)
    proto = RandomConfiguration(0, N, d)
    k = GetKs(proto, K)
    M = len(k)
    v = np.ones(M)

    def configs(first_seed, count):
        return [RandomConfiguration(first_seed + i, N, d) for i in range(count)]

    def sync():
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.synchronize()
        except ImportError:
            pass

    tiles = None
    if args.profile_evals > 0:
        args.profiler_run = True
        args.no_cpu = True
    if args.profiler_run:
        args.no_autotune = True
    if args.tiles:
        from modules_amd import autotune
        vals = [int(t) for t in args.tiles.split(",")]
        seg, rk, rc, fc = vals[:4]
        ppl = vals[4] if len(vals) > 4 else 2
        autotune.record(N, d, M, seg, rk, rc, fc, ppl)
        tiles = {"segment_len": seg, "rho_kernel": rk, "rho_chunk": rc, "force_chunk": fc, "force_ppl": ppl,
                 "source": "--tiles"}
    elif not args.no_autotune:
        from modules_amd import autotune
        tiles = autotune.autotune(N, d, k=k, basis=proto.basis, batch=args.batch, repeats=3)
    solver = BatchSolver(max_slots=max(args.batch, 1), check_interval=args.check_interval, graph=args.graph,
                         precision=args.precision, lanes=args.lanes)

    def run_step(cfgs, clock=None):
        """Submit a batch, admit it (inputs uploaded and resident in HBM), then solve and read the results back.
        Only the solve and the readback are inside `clock` (a list collecting seconds)."""
        ids = [solver.submit(c, k, v, MaxStep=args.maxstep, Emin=args.emin, history=args.history) for c in cfgs]
        solver.step(0)  # admission only: uploads, no evaluation
        sync()
        t0 = time.perf_counter()
        solver.run(chunk=1024)
        res = [solver.result(j) for j in ids]
        sync()
        if clock is not None:
            clock.append(time.perf_counter() - t0)
        for j in ids:
            solver.release(j)
        return res

    if args.profile_evals > 0:
        ids = [solver.submit(c, k, v, MaxStep=args.maxstep, Emin=args.emin, history=args.history)
               for c in configs(0, args.batch)]
        pending = solver.step(args.profile_evals)
        sync()
        if args.copy_probe_mib > 0:
            copy_probe(args.copy_probe_mib)
        print(json.dumps({"profile_evals": args.profile_evals, "batch": args.batch, "N": N, "dim": d, "M": M,
                          "pending": pending, "tiles": tiles}), flush=True)
        return

    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
    warm = [configs(1_000_000 + 1000 * w, args.batch) for w in range(args.warmup)]
    timed = [configs(1000 * s, args.batch) for s in range(args.steps)]
    for cfgs in warm:
        run_step(cfgs)

    solver.set_timing(not args.profiler_run)
    solver.timing(reset=True)
    results = []
    clock = []
    sync()
    t_all = time.perf_counter()
    for cfgs in timed:
        results.extend(run_step(cfgs, clock))
    sync()
    wall_with_upload = time.perf_counter() - t_all
    wall = sum(clock)
    rho_ms, force_ms, n_rho, n_force = solver.timing(reset=True)

    solved = sum(1 for r in results if r["energy"] < args.emin)
    gs_per_s = solved / wall
    dev_evals = sum(r["device_evaluations"] for r in results)
    obj_evals = [r["evaluations"] for r in results]
    lat = np.array([r["t_done"] - r["t_submit"] for r in results])
    pairs = dev_evals * N * M  # (particle, constraint) pairs per contraction, summed over all evaluations
    flops = pairs * (6 * d + 3)  # algorithmic flops per pair per E+F (rho: 2d+1, forces: 4d+2)
    # dominant kernel: the force contraction (k_force_lattice, 4d+2 flops per pair); the rho contraction
    # (k_rho_lattice, 2d+1 flops per pair) is reported beside it
    f_flops = pairs * (4 * d + 2)
    r_flops = pairs * (2 * d + 1)
    achieved = f_flops / (force_ms / 1e3) / 1e12 if force_ms > 0 else 0.0
    achieved_rho = r_flops / (rho_ms / 1e3) / 1e12 if rho_ms > 0 else 0.0
    achieved_ef = flops / ((rho_ms + force_ms) / 1e3) / 1e12 if rho_ms + force_ms > 0 else 0.0

    out = {
        "metric": "stealthy ground states/s at E<1e-18",
        "value": gs_per_s,
        "unit": "ground_states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 phases / f64 sums",
        "data": "synthetic (seeded MT19937 uniform configurations, GetKs constraint set)",
        "config": {"workload": f"batched stealthy relax: {args.batch} seeds x N={N}, d={d}, chi={chi} "
                               f"(M={M}), Emin={args.emin}, 3-pass L-BFGS history {args.history}",
                   "batch": args.batch, "N": N, "dim": d, "chi": chi, "constraints": M, "tiles": tiles,
                   "graph": bool(args.graph), "precision": args.precision, "lanes": args.lanes},
        "solved": solved,
        "solves": len(results),
        "ef_evals_per_s": dev_evals / wall,
        "objective_evals_per_solve_mean": float(np.mean(obj_evals)),
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "pcie_inclusive_value": solved / wall_with_upload,
        "roofline": {
            "bound": "valu",
            "kernel": "k_force_lattice (force contraction, FP64 VALU)",
            "achieved": achieved,
            "peak": FP64_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_VALU_PEAK_TFLOPS,
            "traffic": (_force_traffic(d, N, M, args.batch) or {}).get("bytes_per_launch"),
            "traffic_detail": _force_traffic(d, N, M, args.batch),
            "avg_launch_ms": force_ms / max(n_force, 1),
            "flops_per_pair": 4 * d + 2,
            "rho_kernel": {"kernel": "k_rho_lattice", "achieved": achieved_rho,
                           "frac": achieved_rho / FP64_VALU_PEAK_TFLOPS, "avg_launch_ms": rho_ms / max(n_rho, 1),
                           "flops_per_pair": 2 * d + 1},
            "ef_pair": {"achieved": achieved_ef, "frac": achieved_ef / FP64_VALU_PEAK_TFLOPS,
                        "flops_per_pair": 6 * d + 3},
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(proto, k, v, float(np.mean(obj_evals)), args.cpu_seconds)
    print(json.dumps(out), flush=True)


def _tune(args, cfg, k):
    """Autotuned tiles for one configuration of this shape (unless --no-autotune)."""
    if args.no_autotune:
        return None
    from modules_amd import autotune
    return autotune.autotune(cfg.NumParticle(), cfg.GetDimension(), k=k, basis=cfg.basis, batch=1, repeats=2)


def _force_traffic(d, N, M, batch):
    t = FORCE_TRAFFIC_BYTES.get((d, N, M, batch))
    if t is None:
        return None
    return {"bytes_per_launch": t["read"] + t["write"], "read": t["read"], "write": t["write"],
            "source": "profiles/counters_ef_lbfgs_batched64_3d_N2000_chi045.md"}


def _solve_all(solver, jobs, emin):
    """Admit the submitted jobs (inputs resident in HBM), then run the solver until they finish; returns results
    and the wall seconds of the solve and readback."""
    solver.step(0)
    t0 = time.perf_counter()
    solver.run(chunk=1024)
    wall = time.perf_counter() - t0
    res = [solver.result(j) for j in jobs]
    return res, wall


def workload_single(args):
    """The headline solves through the single-solve drop-in path (host L-BFGS in NLopt order, one GPU E+F pass
    per objective evaluation), on --batch seeds (default 4 here: each solve is sequential)."""
    from modules_amd import GetKs, RandomConfiguration, ShiftedCCPotential, getK
    from modules_amd.optimize import RelaxStructure_NLOPT_Emin
    n_seeds = min(args.batch, 4) if args.batch == 64 else args.batch
    N, d, chi = args.N, args.dim, args.chi
    cfgs = [RandomConfiguration(s, N, d) for s in range(n_seeds)]
    k = GetKs(cfgs[0], getK(chi, 1.0, d))
    pot = ShiftedCCPotential(d)
    for kk in k:
        pot.AddConstraint(kk, [1.0])
    warm = RandomConfiguration(999999, N, d)
    RelaxStructure_NLOPT_Emin(warm, pot, MaxStep=args.maxstep, Emin=args.emin)
    t0 = time.perf_counter()
    res = [RelaxStructure_NLOPT_Emin(c, pot, MaxStep=args.maxstep, Emin=args.emin) for c in cfgs]
    wall = time.perf_counter() - t0
    solved = sum(1 for E, ev, st in res if E < args.emin)
    evals = sum(ev for E, ev, st in res)
    return {"metric": "stealthy ground states/s at E<1e-18", "value": solved / wall, "unit": "ground_states/s",
            "workload": f"single-solve drop-in path: {n_seeds} seeds x N={N}, d={d}, chi={chi} (M={len(k)})",
            "solved": solved, "solves": n_seeds, "wall_s": wall, "objective_evals_per_s": evals / wall,
            "latency_mean_s": wall / n_seeds, "dtype": "f64"}


def workload_mixed256(args):
    """256 independent configurations: N log-uniform in [100, 4000], d in {2, 3}, chi uniform in [0.2, 0.45]
    (one master MT19937 seed draws N, d, chi and the per-configuration seed), all resident in one solver."""
    from modules_amd import GetKs, RandomConfiguration, getK
    from modules_amd.optimize import BatchSolver
    rng = np.random.RandomState(20261203)
    specs = []
    for i in range(256):
        N = int(round(math.exp(rng.uniform(math.log(100), math.log(4000)))))
        d = 2 if rng.uniform() < 0.5 else 3
        chi = float(rng.uniform(0.2, 0.45))
        specs.append((N, d, chi, int(rng.randint(0, 2 ** 31 - 1))))
    solver = BatchSolver(max_slots=256, check_interval=args.check_interval, precision=args.precision)
    jobs, pairs = [], 0
    for N, d, chi, seed in specs:
        cfg = RandomConfiguration(seed, N, d)
        k = GetKs(cfg, getK(chi, 1.0, d))
        jobs.append(solver.submit(cfg, k, np.ones(len(k)), MaxStep=args.maxstep, Emin=args.emin))
        pairs += N * len(k)
    res, wall = _solve_all(solver, jobs, args.emin)
    solved = sum(1 for r in res if r["energy"] < args.emin)
    lat = np.array([r["t_done"] - r["t_submit"] for r in res])
    return {"metric": "stealthy ground states/s at E<1e-18", "value": solved / wall, "unit": "ground_states/s",
            "workload": "256 configurations, N log-uniform [100,4000], d in {2,3}, chi U[0.2,0.45]",
            "solved": solved, "solves": len(res), "wall_s": wall,
            "ef_evals_per_s": sum(r["device_evaluations"] for r in res) / wall,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "pairs_per_contraction_total": pairs, "dtype": "f64"}


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


def workload_mixed2d(args):
    """2D N=20000 at density 1, chi=0.4: the mixed-precision error at the start configuration (E and forces of
    the mixed path against the FP64 path, which the tests pin to the oracle at 1e-12), then the same solve in
    FP64 and in mixed precision (FP32 phases and force terms, FP64 sums, FP64 polish once E < 1e-8)."""
    from modules_amd import GetKs, RandomConfiguration, ShiftedCCPotential, getK
    from modules_amd.optimize import BatchSolver
    cfg = RandomConfiguration(4, 20000, 2)
    k = GetKs(cfg, getK(0.4, 1.0, 2))
    v = np.ones(len(k))
    tiles = _tune(args, cfg, k)
    pot = ShiftedCCPotential(2)
    for kk in k:
        pot.AddConstraint(kk, [1.0])
    pot.SetConfiguration(cfg)
    E64, F64 = pot.EnergyAndAllForce()
    pot.SetPrecision("mixed")
    pot.SetConfiguration(cfg)
    E32, F32 = pot.EnergyAndAllForce()
    out = {"metric": "stealthy ground states/s at E<1e-18", "unit": "ground_states/s",
           "workload": f"2D N=20000 chi=0.4 (M={len(k)}): mixed precision against FP64", "tiles": tiles,
           "start_config_error": {"dE_over_E": abs(E32 - E64) / abs(E64),
                                  "max_dF_over_max_F": float(np.max(np.abs(F32 - F64)) / np.max(np.abs(F64)))}}
    for prec in ("fp64", "mixed"):
        solver = BatchSolver(max_slots=1, check_interval=args.check_interval, precision=prec)
        job = solver.submit(cfg, k, v, MaxStep=args.maxstep, Emin=args.emin)
        res, wall = _solve_all(solver, [job], args.emin)
        r = res[0]
        out[prec] = {"value": (1.0 if r["energy"] < args.emin else 0.0) / wall, "wall_s": wall,
                     "energy": r["energy"], "objective_evals": r["evaluations"],
                     "device_evals": r["device_evaluations"], "switch_eval": r["switch_eval"],
                     "ef_evals_per_s": r["device_evaluations"] / wall}
    out["value"] = out["mixed"]["value"]
    out["speedup_mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"] if out["fp64"]["value"] else None
    out["dtype"] = "f32 phases / f64 sums"
    return out


def workload_serve(args):
    """Poisson request stream (rate --rate/s, --requests requests): N lognormal (median 1000, sigma_ln 0.6,
    clipped to [100, 50000]), d in {2,3}, chi U[0.1, 0.49], soft core e0 (1 - r/sigma)^2 with e0 = 1 and
    sigma = 0.4 density^(-1/d); continuous batching on one solver."""
    from modules_amd.serving import make_stream, serve
    reqs = make_stream(args.requests, args.rate)
    reqs, wall = serve(reqs, Emin=args.emin, MaxStep=args.maxstep, max_slots=64)
    lat = np.array([r.done - r.arrival for r in reqs])
    solved = sum(1 for r in reqs if r.result.get("energy", 1.0) < args.emin)
    return {"metric": "stealthy ground states/s at E<1e-18", "value": solved / wall, "unit": "ground_states/s",
            "workload": f"serving: {args.requests} Poisson requests at {args.rate}/s, soft core e0=1 sigma=0.4",
            "solved": solved, "requests": len(reqs), "wall_s": wall,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "N_median": float(np.median([r.N for r in reqs])), "dtype": "f64"}


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

    out = {"metric": "stealthy ground states/s at E<1e-18", "unit": "ground_states/s", "dtype": "f64"}
    cfg = RandomConfiguration(4, 20000, 2)
    k = GetKs(cfg, getK(0.4, 1.0, 2))
    tiles = _tune(args, cfg, k)
    ms_d, E_d, F_d, _ = ef_ms(cfg, k, "direct", 20)
    ms_n, E_n, F_n, used = ef_ms(cfg, k, "nufft", 20)
    out["N20000"] = {"M": len(k), "tiles": tiles, "backend_used": used,
                     "ef_ms_direct": ms_d, "ef_ms_nufft": ms_n,
                     "nufft_vs_direct_start": {"dE_over_E": abs(E_n - E_d) / abs(E_d),
                                               "max_dF_over_max_F": float(np.max(np.abs(F_n - F_d)) / np.max(np.abs(F_d)))},
                     "direct": solve(cfg, k, "direct"), "nufft": solve(cfg, k, "nufft")}
    cfg = RandomConfiguration(5, 100000, 2)
    k = GetKs(cfg, getK(0.4, 1.0, 2))
    ms_d, E_d, F_d, _ = ef_ms(cfg, k, "direct", 2)
    ms_n, E_n, F_n, used = ef_ms(cfg, k, "nufft", 10)
    r = solve(cfg, k, "nufft")
    out["N100000"] = {"M": len(k), "backend_used": used, "ef_ms_direct": ms_d, "ef_ms_nufft": ms_n,
                      "nufft_vs_direct_start": {"dE_over_E": abs(E_n - E_d) / abs(E_d),
                                                "max_dF_over_max_F": float(np.max(np.abs(F_n - F_d)) / np.max(np.abs(F_d)))},
                      "nufft": r, "direct_solve_estimate_s": ms_d * r["device_evals"] / 1000}
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


OTHER_WORKLOADS = {"single": workload_single, "mixed256": workload_mixed256, "large2d": workload_large2d,
                   "mixed2d": workload_mixed2d, "serve": workload_serve, "five": workload_five,
                   "nufft2d": workload_nufft2d, "targets": workload_targets, "md": workload_md}


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


def cpu_baseline(cfg, k, v, evals_per_gs, seconds):
    """The oracle's E+F (the reference direct path restated in C with OpenMP, libm sin/cos per pair) timed on
    this host for a bounded number of evaluations of the same configuration; ground states/s = E+F/s divided
    by the objective evaluations one ground state takes (the GPU run's mean; the optimizer is the same).  Also
    a 1-thread sample and the reference's published sanity anchor (2D N=500 chi=0.2996 on 2 threads:
    4.0535 ms per E+F, docs/finufft/reports/one-component-benchmark.md:37)."""
    os.environ.setdefault("OMP_WAIT_POLICY", "ACTIVE")
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from modules_amd import GetKs, RandomConfiguration, getK
    from oracle import pyoracle as po

    threads = min(16, os.cpu_count() or 1)
    t = _time_ef(po, cfg, k, v, threads, seconds)
    ef_per_s = len(t) / t.sum()
    t1 = _time_ef(po, cfg, k, v, 1, min(3.0, seconds / 5))
    acfg = RandomConfiguration(0, 500, 2)
    ak = GetKs(acfg, getK(0.2996, 1.0, 2))
    ta = _time_ef(po, acfg, ak, np.ones(len(ak)), 2, 1.0, min_n=30)
    return {"value": ef_per_s / evals_per_gs, "unit": "ground_states/s", "cores": threads, "kind": "port",
            "sample": f"{len(t)} oracle E+F evaluations of one N={cfg.NumParticle()} d={cfg.GetDimension()} "
                      f"configuration on {threads} threads in {t.sum():.1f} s ({ef_per_s:.3g} E+F/s, median "
                      f"{1e3 * np.median(t):.1f} ms) divided by {evals_per_gs:.0f} evaluations per ground state",
            "ef_evals_per_s": ef_per_s, "median_ms_per_ef": 1e3 * float(np.median(t)),
            "one_thread_median_ms_per_ef": 1e3 * float(np.median(t1)),
            "anchor_2d_N500_chi0.2996_2threads": {"median_ms_per_ef": 1e3 * float(np.median(ta)), "samples": len(ta),
                                                  "constraints": len(ak), "published_ms": 4.0535},
            "cpu_model": _cpu_model(), "host_logical_cpus": os.cpu_count(),
            "omp": {"OMP_WAIT_POLICY": os.environ.get("OMP_WAIT_POLICY"),
                    "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND")}}


if __name__ == "__main__":
    main()
