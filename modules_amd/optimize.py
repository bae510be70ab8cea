"""Structure optimisation on the GPU engine, mirroring potential/StructureOptimization.{h,cpp}.

  RelaxStructure_NLOPT / RelaxStructure_NLOPT_Emin   StructureOptimization.cpp:1046-1084 (Switch 0: particles move)
  LuksanPLIS                                          potential/LuksanPLIS.h:35-36
  CollectiveCoordinateMultiRun                        apps/CCOptimization/main.cpp:1811-1872 (the solve loop)
  BatchSolver                                         many RelaxStructure solves resident on the device at once
                                                      (continuous batching; no counterpart in the reference)
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import check, dp, lib
from .potential import Configuration, ShiftedCCPotential

MaxEnergy = 1e10  # core/etc.h:8
StructureOptimization_LBFGSHistory = 40  # potential/StructureOptimization.cpp:17
# potential/StructureOptimization.cpp:18 (0 nlopt, 1 ring, 3 fast): the selector is kept for the drivers' commands;
# every value runs the one device L-BFGS (Luksan's limited-memory BFGS, the algorithm of the nlopt choice)
StructureOptimization_LBFGSImplementation = 0


def _p(a):
    return a.ctypes.data_as(dp)


def LuksanPLIS(fun, x0, stopval=-np.inf, maxeval=0, mf=40):
    """Minimise fun(x) -> (f, grad) with PLIS in NLopt's arithmetic.  Returns (x, minf, status, evaluations)."""
    x = np.ascontiguousarray(x0, dtype=np.float64).copy()
    n = len(x)

    @_lib.OBJECTIVE
    def obj(nn, xp, gp, data):
        xv = np.ctypeslib.as_array(xp, shape=(nn,))
        f, g = fun(xv.copy())
        np.ctypeslib.as_array(gp, shape=(nn,))[:] = g
        return float(f)

    minf = C.c_double(0)
    ev = C.c_int(0)
    st = lib().ccgpu_lbfgs_minimize(n, C.cast(obj, C.c_void_p), None, _p(x), C.byref(minf), stopval, maxeval, mf,
                                    C.byref(ev))
    return x, minf.value, st, ev.value


def _check_switch(Pressure, Switch, MinDistance):
    if Switch not in (0, 1, 2):
        raise ValueError("Switch must be 0 (particles), 1 (cell) or 2 (both)")
    if MinDistance > 1e-4:
        raise NotImplementedError("the MinDistance neighbour check (MinDistance > 1e-4) is not on this path")


def RelaxStructure_NLOPT_Emin(List: Configuration, pot: ShiftedCCPotential, Pressure=0.0, Switch=0, MinDistance=0.0,
                              MaxStep=10000, Emin=-MaxEnergy, history=None):
    """Relax List in place with three L-BFGS passes (rescale 10000, 1, 0.01), stopping once E < Emin.
    The L-BFGS runs on the host, each objective evaluation is one GPU E+F pass.  Returns (E, evaluations, status);
    with a Pressure or a moving cell (Switch 1: cell, 2: both; variable-cell potentials) the objective is the
    enthalpy E + P V and the first value returned is the final enthalpy."""
    _check_switch(Pressure, Switch, MinDistance)
    history = history or StructureOptimization_LBFGSHistory
    pot.SetConfiguration(List)
    if Switch != 0 or Pressure != 0:
        basis = np.ascontiguousarray(List.basis, dtype=np.float64).copy()
        rel = np.ascontiguousarray(List.relative, dtype=np.float64).copy()
        H = C.c_double(0)
        ev = C.c_int(0)
        st = C.c_int(0)
        check(lib().ccgpu_potential_relax_cell(pot._h, _p(basis), _p(rel), float(Pressure), int(Switch),
                                               float(MinDistance), int(MaxStep), float(Emin), int(history),
                                               C.byref(H), C.byref(ev), C.byref(st)))
        List.basis = basis.reshape(List.basis.shape)
        List.relative = rel.reshape(List.relative.shape)
        pot.RhoValid = False
        if hasattr(pot, "_pull"):
            pot._config = List
            pot._pull()
        return H.value, ev.value, st.value
    rel = np.ascontiguousarray(List.relative, dtype=np.float64).copy()
    E = C.c_double(0)
    ev = C.c_int(0)
    st = C.c_int(0)
    check(lib().ccgpu_potential_relax(pot._h, _p(rel), int(MaxStep), float(Emin), int(history), C.byref(E),
                                      C.byref(ev), C.byref(st)))
    List.relative = rel.reshape(List.relative.shape)
    pot.RhoValid = False
    return E.value, ev.value, st.value


def RelaxStructure_NLOPT(List, pot, Pressure=0.0, Switch=0, MinDistance=0.0, MaxStep=10000, history=None):
    return RelaxStructure_NLOPT_Emin(List, pot, Pressure, Switch, MinDistance, MaxStep, -MaxEnergy, history)


def _potential_terms(pot):
    """Constraint set, shift and soft-core parameters of a potential, as the solver takes them."""
    d = pot.Dimension
    k = np.array([c.k[:d] for c in pot.constraints], dtype=np.float64).reshape(-1, d)
    V = np.array([c.V for c in pot.constraints], dtype=np.float64)
    e0, sigma = getattr(pot, "_softcore", (0.0, 0.0))
    return k, V, float(pot.Shift), float(e0), float(sigma)


def _multi_terms(pot):
    """Species lists, constraints, radii and scales of a MultiStealthyPotential, as the solver takes them."""
    d = pot.Dimension
    pot._require()
    lists = [sp.idx for sp in pot.species_constraints]
    idx_off = np.zeros(len(lists) + 1, dtype=np.int64)
    k_off = np.zeros(len(lists) + 1, dtype=np.int64)
    ks, vs = [], []
    for q, sp in enumerate(pot.species_constraints):
        idx_off[q + 1] = idx_off[q] + len(lists[q])
        cons = sp.CCPotential.constraints
        k_off[q + 1] = k_off[q] + len(cons)
        ks.append(np.array([c.k[:d] for c in cons], dtype=np.float64).reshape(-1, d))
        vs.append(np.array([c.V for c in cons], dtype=np.float64))
    return {"idx_off": idx_off, "k_off": k_off,
            "idx": np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.int64) for x in lists] +
                                                       [np.zeros(1, dtype=np.int64)])),
            "k": np.ascontiguousarray(np.concatenate(ks + [np.zeros((1, d))])),
            "V": np.ascontiguousarray(np.concatenate(vs + [np.zeros(1)])),
            "radii": np.ascontiguousarray(pot.GetRadii(), dtype=np.float64), "rc": float(pot.GetRc()),
            "u0": float(pot.u0), "e0": float(pot.e0) if pot.WithRepulsion else 0.0}


def CollectiveCoordinateMultiRun(pConfig: Configuration, pPotential, gen, Prefix="",
                                 SampleNumber=1, TimeLimit=None, AlgorithmName="LBFGS", E_max=1e-14, max_num=10000,
                                 onAccepted=None, max_slots=64, verbose=False):
    """CollectiveCoordinateMultiRun (apps/CCOptimization/main.cpp:1811-1872) on the batched device solver.

    Sample i starts from pConfig for i = 0 and from the next N*d draws of `gen` (a RandomGenerator) after
    sample i-1, exactly the starts of the reference's serial loop; each is relaxed with RelaxStructure_NLOPT
    (MaxStep = max_num), its energy is re-evaluated with pPotential and it is accepted when E < E_max.  All
    samples are submitted at once and relax side by side.  pPotential is a ShiftedCCPotential (optionally with the
    soft core) or a MultiStealthyPotential (GetMultiCCO's ground mode); a mixture keeps its species and radii.  Results are handed out in sample order, so
    onAccepted(config, success_index) sees the reference's order.  TimeLimit (absolute time.time() seconds, as
    the reference's time_t) is checked after each sample; reaching it ends the run without re-randomizing
    pConfig.  Returns {"init": [...], "relaxed": [...], "success": [...], "energies": [...]}; with a non-empty
    Prefix the same configurations are also appended to the ConfigurationPack files Prefix (relaxed),
    Prefix_InitConfig and Prefix_Success, as the reference writes them, and onAccepted gets the success pack's
    index.
    """
    if AlgorithmName != "LBFGS":
        raise NotImplementedError(f"CollectiveCoordinateMultiRun: only LBFGS is on this path (got {AlgorithmName})")
    N, d = pConfig.NumParticle(), pConfig.GetDimension()
    multi = hasattr(pPotential, "species_constraints")  # MultiStealthyPotential (GetMultiCCO, main.cpp:1565-1590)
    if not multi:
        k, V, shift, e0, sigma = _potential_terms(pPotential)
    starts = [pConfig.copy()]
    pos0 = getattr(gen, "_pos", None)
    for _ in range(1, SampleNumber):
        starts.append(Configuration(pConfig.basis, gen.RandomDoubles(N * d).reshape(N, d)))
    out = {"init": [], "relaxed": [], "success": [], "energies": []}
    packs = None
    if Prefix:
        # Prefix (relaxed), Prefix_Success and Prefix_InitConfig ConfigurationPacks (main.cpp:1816-1818)
        from .formats import ConfigurationPack
        packs = (ConfigurationPack(Prefix), ConfigurationPack(Prefix + "_Success"),
                 ConfigurationPack(Prefix + "_InitConfig"))
    if SampleNumber <= 0:
        return out
    solver = BatchSolver(max_slots=max(1, min(max_slots, SampleNumber)))
    if multi:
        ids = [solver.submit_multi(c, pPotential, MaxStep=max_num) for c in starts]
    elif getattr(pPotential, "_S2", False):
        if hasattr(pPotential, "_pull"):
            raise NotImplementedError("CollectiveCoordinateMultiRun: variable-cell potentials relax through "
                                      "RelaxStructure_NLOPT")
        S0 = np.asarray(pPotential.S0, dtype=np.float64)
        ids = [solver.submit_s2(c, k, V, S0, shift=shift, epsilon=e0, sigma=sigma, MaxStep=max_num) for c in starts]
    else:
        ids = [solver.submit(c, k, V, shift=shift, epsilon=e0, sigma=sigma, MaxStep=max_num) for c in starts]
    nxt = 0
    while nxt < SampleNumber:
        pending = solver.step(256)
        while nxt < SampleNumber:
            r = solver.result(ids[nxt])
            if r["state"] == BatchSolver.FAILED:
                raise RuntimeError(f"CollectiveCoordinateMultiRun: sample {nxt} failed")
            if r["state"] != BatchSolver.DONE:
                break
            init = starts[nxt]
            pPotential.SetConfiguration(init)
            E_init = pPotential.Energy()
            result = Configuration(pConfig.basis, r["relative"])
            pPotential.SetConfiguration(result)
            E = pPotential.Energy()
            if verbose:
                print(f"generating config{nxt}\n \tE_init={E_init} \n \tE_relax={E} ")
            out["init"].append(init)
            out["relaxed"].append(result)
            out["energies"].append(E)
            if packs:  # the reference's order: start, relaxed result, then the success pack
                packs[2].AddConfig(init)
                packs[0].AddConfig(result)
            if E < E_max:
                out["success"].append(result)
                if packs:
                    packs[1].AddConfig(result)
                if onAccepted is not None:
                    onAccepted(result, len(out["success"]) - 1 if not packs else packs[1].NumConfig() - 1)
            solver.release(ids[nxt])
            nxt += 1
            if TimeLimit is not None and time.time() > TimeLimit:
                # the serial loop returns here before re-randomizing: pConfig still holds this sample's start and
                # the generator has given out only the draws of the samples started so far
                pConfig.relative = starts[nxt - 1].relative.copy()
                if pos0 is not None:
                    gen._pos = pos0 + (nxt - 1) * N * d
                return out
        if pending == 0 and nxt < SampleNumber:
            raise RuntimeError("CollectiveCoordinateMultiRun: solver drained with samples unfinished")
    pConfig.relative = _wrap_rel(gen.RandomDoubles(N * d).reshape(N, d))
    return out


def _wrap_rel(s):
    y = s - np.floor(s)
    return y - np.floor(y)


class BatchSolver:
    """Device-resident solver: jobs (RelaxStructure_NLOPT[_Emin] requests) are admitted into slots by HBM budget
    and advance together, one batched E+F launch and one L-BFGS step launch per evaluation."""

    QUEUED, RESIDENT, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, max_slots=256, hbm_budget=200 << 30, check_interval=8, graph=False, precision="fp64",
                 polish=1e-8, lanes=2, rho_backend="direct", nufft_eps=1e-12):
        self._h = C.c_void_p()
        check(lib().ccgpu_solver_create(int(max_slots), int(hbm_budget), C.byref(self._h)))
        check(lib().ccgpu_solver_set_check_interval(self._h, int(check_interval)))
        check(lib().ccgpu_solver_set_graph(self._h, 1 if graph else 0))
        check(lib().ccgpu_solver_set_lanes(self._h, int(lanes)))
        self.set_precision(precision, polish)
        self.set_rho_backend(rho_backend, nufft_eps)
        self._shapes = {}

    def set_lanes(self, lanes):
        """Solver streams (1 or 2) for jobs admitted from now on; only while no job is resident."""
        check(lib().ccgpu_solver_set_lanes(self._h, int(lanes)))

    def set_graph(self, on=True):
        """Replay the optimizer steps between two host checks from a captured hipGraph (True) or launch them on
        the stream (False)."""
        check(lib().ccgpu_solver_set_graph(self._h, 1 if on else 0))

    def set_precision(self, precision="fp64", polish=1e-8):
        """Precision of jobs submitted from now on: "fp64", or "mixed" (FP32 phases and force terms, FP64
        sums) until E < polish or the end of the first pass, FP64 afterwards."""
        from .potential import _precision_code
        check(lib().ccgpu_solver_set_precision(self._h, _precision_code(precision), float(polish)))

    def set_rho_backend(self, backend="direct", eps=1e-12):
        """Collective-density backend of single-species jobs submitted from now on: "direct", "nufft" (eligible
        lattice constraint sets the cost gate picks) or "nufft_always" (every eligible set); tolerance eps."""
        codes = {"direct": 0, "nufft": 1, "nufft_always": 2}
        if backend not in codes:
            raise ValueError('backend must be "direct", "nufft" or "nufft_always"')
        check(lib().ccgpu_solver_set_rho_backend(self._h, codes[backend], float(eps)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            lib().ccgpu_solver_destroy(h)
            self._h = None

    def submit(self, config: Configuration, k, V, shift=0.0, epsilon=0.0, sigma=0.0, MaxStep=10000, Emin=-MaxEnergy,
               history=None):
        history = history or StructureOptimization_LBFGSHistory
        d = config.GetDimension()
        rel = np.ascontiguousarray(config.relative, dtype=np.float64)
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, d)
        V = np.ascontiguousarray(V, dtype=np.float64)
        job = C.c_int64(0)
        check(lib().ccgpu_solver_submit(self._h, d, _p(np.ascontiguousarray(config.basis)), rel.shape[0], _p(rel),
                                        k.shape[0], _p(k), _p(V), shift, epsilon, sigma, int(MaxStep), float(Emin),
                                        int(history), C.byref(job)))
        self._shapes[job.value] = rel.shape
        return job.value

    def submit_s2(self, config: Configuration, k, V, S0, shift=0.0, epsilon=0.0, sigma=0.0, MaxStep=10000,
                  Emin=-MaxEnergy, history=None):
        """A relax request with the target-S(k) potential (ShiftedCCPotential_S2; soft core as
        RepulsiveCCPotential with S0 != 0): one target S0 per constraint."""
        history = history or StructureOptimization_LBFGSHistory
        d = config.GetDimension()
        rel = np.ascontiguousarray(config.relative, dtype=np.float64)
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(-1, d)
        V = np.ascontiguousarray(V, dtype=np.float64)
        S0 = np.ascontiguousarray(S0, dtype=np.float64)
        if S0.shape != V.shape:
            raise ValueError("one target S0 per constraint")
        job = C.c_int64(0)
        check(lib().ccgpu_solver_submit_s2(self._h, d, _p(np.ascontiguousarray(config.basis)), rel.shape[0],
                                           _p(rel), k.shape[0], _p(k), _p(V), _p(S0 if len(S0) else np.zeros(1)),
                                           shift, epsilon, sigma, int(MaxStep), float(Emin), int(history),
                                           C.byref(job)))
        self._shapes[job.value] = rel.shape
        return job.value

    def submit_multi(self, config: Configuration, pot, MaxStep=10000, Emin=-MaxEnergy,
                     history=None):
        """A relax request with a MultiStealthyPotential `pot` (its species lists, constraints, radii, u0, e0 and
        contact cutoff) from the positions of `config` (same particles as pot's configuration)."""
        history = history or StructureOptimization_LBFGSHistory
        d = config.GetDimension()
        rel = np.ascontiguousarray(config.relative, dtype=np.float64)
        N = rel.shape[0]
        terms = _multi_terms(pot)
        if len(terms["radii"]) != N:
            raise ValueError("configuration and potential hold different particle numbers")
        i64 = C.POINTER(C.c_int64)
        job = C.c_int64(0)
        check(lib().ccgpu_solver_submit_multi(
            self._h, d, _p(np.ascontiguousarray(config.basis)), N, _p(rel), len(terms["idx_off"]) - 1,
            terms["idx_off"].ctypes.data_as(i64), terms["idx"].ctypes.data_as(i64), terms["k_off"].ctypes.data_as(i64),
            _p(terms["k"]), _p(terms["V"]), _p(terms["radii"]), terms["rc"], terms["u0"], terms["e0"], int(MaxStep),
            float(Emin), int(history), C.byref(job)))
        self._shapes[job.value] = rel.shape
        return job.value

    def step(self, evaluations=64):
        pending = C.c_int(0)
        check(lib().ccgpu_solver_step(self._h, int(evaluations), C.byref(pending)))
        return pending.value

    def run(self, chunk=256, timeout=None):
        t0 = time.time()
        while self.step(chunk) > 0:
            if timeout is not None and time.time() - t0 > timeout:
                raise TimeoutError("BatchSolver.run: timeout")

    def set_time_limit(self, seconds):
        """Wall-clock limit (seconds from admission) of jobs submitted from now on; 0: none.  A job past its limit
        is stopped and returned with status 6 (NLOPT_MAXTIME_REACHED) and its last evaluated configuration."""
        check(lib().ccgpu_solver_set_time_limit(self._h, float(seconds)))

    def result(self, job):
        """State, energy, counts, times and relaxed configuration of a job.  A failed job (state FAILED) also
        carries "error" (the message) and "code" (CCGPU_ERR_NAN for a NaN energy or non-finite gradient,
        the reference's NotANumberFound; CCGPU_ERR_MEMORY or CCGPU_ERR_ARG when admission failed)."""
        state = C.c_int(0)
        E = C.c_double(0)
        ev = C.c_int(0)
        dev = C.c_int(0)
        st = C.c_int(0)
        t = np.zeros(3)
        rel = np.zeros(self._shapes[job])
        code = lib().ccgpu_solver_result(self._h, int(job), C.byref(state), C.byref(E), C.byref(ev), C.byref(dev),
                                         C.byref(st), _p(t), _p(rel))
        out = {"state": state.value, "energy": E.value, "evaluations": ev.value, "device_evaluations": dev.value,
               "status": st.value, "t_submit": t[0], "t_admit": t[1], "t_done": t[2], "relative": rel}
        if code != 0:
            if state.value != self.FAILED:
                check(code)
            out["code"] = code
            out["error"] = lib().ccgpu_last_error().decode()
            out["switch_eval"] = 0
            return out
        sw = C.c_int(0)
        check(lib().ccgpu_solver_switch_eval(self._h, int(job), C.byref(sw)))
        out["switch_eval"] = sw.value
        return out

    def release(self, job):
        check(lib().ccgpu_solver_release(self._h, int(job)))
        self._shapes.pop(job, None)

    def set_timing(self, on=True):
        check(lib().ccgpu_solver_set_timing(self._h, 1 if on else 0))

    def timing(self, reset=False):
        """Accumulated (rho_ms, force_ms, rho_launches, force_launches) of the contraction kernels."""
        a = C.c_double(0)
        b = C.c_double(0)
        na = C.c_int64(0)
        nb = C.c_int64(0)
        check(lib().ccgpu_solver_timing(self._h, C.byref(a), C.byref(b), C.byref(na), C.byref(nb), 1 if reset else 0))
        return a.value, b.value, na.value, nb.value

    def job_bytes(self, job):
        """(estimate charged against the HBM budget, device bytes held while resident) of a job."""
        e = C.c_int64(0)
        r = C.c_int64(0)
        check(lib().ccgpu_solver_job_bytes(self._h, int(job), C.byref(e), C.byref(r)))
        return e.value, r.value

    def launch_stats(self):
        """Bursts of optimizer steps launched so far: {"graph": replayed hipGraphs, "stream": stream-launched}."""
        g = C.c_int64(0)
        st = C.c_int64(0)
        check(lib().ccgpu_solver_launch_stats(self._h, C.byref(g), C.byref(st)))
        return {"graph": g.value, "stream": st.value}

    def stats(self):
        r = C.c_int(0)
        q = C.c_int(0)
        b = C.c_int64(0)
        check(lib().ccgpu_solver_stats(self._h, C.byref(r), C.byref(q), C.byref(b)))
        return {"resident": r.value, "queued": q.value, "bytes_in_use": b.value}
