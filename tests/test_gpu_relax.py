"""GPU stealthy ground-state solves: the single-solve drop-in path (host L-BFGS, GPU E+F) and the batched
device-resident solver, checked against the oracle relax driven by the reference LuksanPLIS."""
import numpy as np
import pytest

from modules_amd import GetKs, RandomConfiguration, RepulsiveCCPotential, ShiftedCCPotential, getK
from modules_amd.optimize import BatchSolver, LuksanPLIS, RelaxStructure_NLOPT_Emin
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

EMIN = 1e-18


def stealthy(N, d, chi, seed):
    cfg = RandomConfiguration(seed, N, d)
    k = GetKs(cfg, getK(chi, 1.0, d))
    return cfg, k, np.ones(len(k))


def test_single_solve_reaches_ground_state_like_reference():
    cfg, k, v = stealthy(200, 2, 0.3, 0)
    pot = ShiftedCCPotential(2)
    for kk in k:
        pot.AddConstraint(kk, [1.0])
    E, evals, st = RelaxStructure_NLOPT_Emin(cfg, pot, MaxStep=10000, Emin=EMIN)
    Eo, so, evo, sto = po.relax_emin(po.make_config(0, 200, 2), cfg.basis, k, v, Emin=EMIN)
    assert E < EMIN and Eo < EMIN
    assert st == 2 and sto == 2
    assert 0.5 * evo < evals < 2.0 * evo
    # the relaxed configuration is a ground state for the oracle too
    Ef, Ff = po.total_energy_forces(cfg.relative, cfg.basis, k, v)
    assert Ef < 10 * EMIN


def test_objective_matches_oracle_every_100th_evaluation():
    # configuration 2 style check (3D, chi 0.45) at a reduced size: E and gradient of the GPU objective
    # against the oracle at every 100th evaluation of an L-BFGS pass
    N, d = 400, 3
    cfg, k, v = stealthy(N, d, 0.45, 3)
    pot = ShiftedCCPotential(d)
    for kk in k:
        pot.AddConstraint(kk, [1.0])
    a = cfg.basis
    vol = po.volume(a)
    escale = len(k) * N * N / vol
    fscale = 4 * len(k) * N * np.max(np.linalg.norm(k, axis=1)) / vol
    count = [0]
    worst = [0.0, 0.0]

    def fun(x):
        s = x.reshape(N, d) % 1.0
        from modules_amd import Configuration
        c = Configuration(a, s)
        pot.SetConfiguration(c)
        E, F = pot.EnergyAndAllForce()
        if count[0] % 100 == 0:
            Eo, Fo = po.total_energy_forces(c.relative, a, k, v)
            worst[0] = max(worst[0], abs(E - Eo) / escale)
            worst[1] = max(worst[1], np.max(np.abs(F - Fo)) / fscale)
        count[0] += 1
        return E, -(F @ a.T).ravel()

    LuksanPLIS(fun, cfg.relative.ravel(), stopval=EMIN, maxeval=600)
    assert count[0] >= 200
    assert worst[0] <= 1e-12 and worst[1] <= 1e-12


def test_batch_solver_mixed_sizes_all_reach_ground_state():
    solver = BatchSolver(max_slots=16)
    jobs = []
    specs = [(150, 2, 0.3, 1), (300, 3, 0.4, 2), (500, 2, 0.2, 3), (120, 3, 0.45, 4), (800, 2, 0.35, 5)]
    for N, d, chi, seed in specs:
        cfg, k, v = stealthy(N, d, chi, seed)
        jobs.append((solver.submit(cfg, k, v, MaxStep=10000, Emin=EMIN), cfg, k, v))
    solver.run(chunk=128, timeout=300)
    for job, cfg, k, v in jobs:
        r = solver.result(job)
        assert r["state"] == BatchSolver.DONE
        assert r["energy"] < EMIN
        Eo, _ = po.total_energy_forces(r["relative"], cfg.basis, k, v)
        assert Eo < 10 * EMIN
        assert 10 < r["evaluations"] < 20000


def test_batch_result_is_the_same_alone_and_in_the_mix():
    cfg, k, v = stealthy(400, 3, 0.4, 11)
    alone = BatchSolver(max_slots=4)
    j0 = alone.submit(cfg, k, v, Emin=EMIN)
    alone.run(timeout=300)
    ra = alone.result(j0)
    mix = BatchSolver(max_slots=8)
    others = [stealthy(N, d, chi, s) for N, d, chi, s in ((250, 2, 0.3, 21), (600, 3, 0.25, 22), (90, 2, 0.45, 23))]
    mix.submit(*others[0], Emin=EMIN)
    j1 = mix.submit(cfg, k, v, Emin=EMIN)
    for o in others[1:]:
        mix.submit(*o, Emin=EMIN)
    mix.run(timeout=300)
    rm = mix.result(j1)
    assert ra["energy"] == rm["energy"]
    assert ra["evaluations"] == rm["evaluations"]
    assert np.array_equal(ra["relative"], rm["relative"])


def test_batch_admission_by_budget_and_release():
    solver = BatchSolver(max_slots=2, hbm_budget=64 << 20)
    ids = [solver.submit(*stealthy(200, 2, 0.3, s), Emin=EMIN) for s in range(5)]
    st = solver.stats()
    assert st["resident"] == 0 and st["queued"] == 5
    solver.step(1)
    assert solver.stats()["resident"] <= 2
    solver.run(timeout=300)
    for j in ids:
        assert solver.result(j)["energy"] < EMIN
        solver.release(j)
    assert solver.stats()["bytes_in_use"] == 0


def test_softcore_energy_forces_match_oracle():
    for d, N, sigma in ((2, 300, 0.4), (3, 250, 0.4 * 1.0)):
        cfg, k, v = stealthy(N, d, 0.3, 7)
        pot = RepulsiveCCPotential(d, 2.0, sigma)
        for kk in k:
            pot.AddConstraint(kk, [1.0])
        pot.SetConfiguration(cfg)
        E, F = pot.EnergyAndAllForce()
        Eo, Fo = po.total_energy_forces(cfg.relative, cfg.basis, k, v, sigma=sigma, e0=2.0)
        assert abs(E - Eo) <= 1e-12 * max(1.0, abs(Eo))
        assert np.max(np.abs(F - Fo)) <= 1e-11 * max(1.0, np.max(np.abs(Fo)))


def test_graph_replay_matches_stream_launches():
    # the same solve with the optimizer steps replayed from a captured hipGraph (many replays, host checks and
    # harvests in between) and with plain stream launches: bitwise the same result
    specs = [stealthy(300, 2, 0.3, 31), stealthy(200, 3, 0.4, 32)]
    out = []
    for graph in (False, True):
        solver = BatchSolver(max_slots=4, check_interval=8, graph=graph)
        ids = [solver.submit(*sp, Emin=EMIN) for sp in specs]
        solver.run(timeout=300)
        out.append([solver.result(j) for j in ids])
    for a, b in zip(*out):
        assert a["energy"] == b["energy"] and a["evaluations"] == b["evaluations"]
        assert np.array_equal(a["relative"], b["relative"])


def test_multirun_starts_acceptance_and_generator_state():
    # CollectiveCoordinateMultiRun (apps/CCOptimization/main.cpp:1811-1872): sample i starts from the i-th block
    # of N*d draws of the generator, every relaxed sample is re-evaluated and accepted below E_max, acceptance is
    # reported in sample order and pConfig ends on the next block of draws
    from modules_amd.optimize import CollectiveCoordinateMultiRun
    from modules_amd.potential import RandomGenerator
    N, d, S = 120, 2, 5
    gen = RandomGenerator(42)
    cfg = RandomConfiguration(gen, N, d)
    k = GetKs(cfg, getK(0.3, 1.0, d))
    pot = ShiftedCCPotential(d)
    for kk in k:
        pot.AddConstraint(kk, [1.0])
    seen = []
    out = CollectiveCoordinateMultiRun(cfg, pot, gen, SampleNumber=S, E_max=1e-14, max_num=10000,
                                       onAccepted=lambda c, i: seen.append(i))
    stream = RandomConfiguration(42, (S + 1) * N, d).relative
    for i, c in enumerate(out["init"]):
        assert np.array_equal(c.relative, stream[i * N:(i + 1) * N])
    assert np.array_equal(cfg.relative, stream[S * N:])
    assert len(out["relaxed"]) == S
    assert seen == list(range(len(out["success"])))
    for c, E in zip(out["relaxed"], out["energies"]):
        Eo, _ = po.total_energy_forces(c.relative, c.basis, k, np.ones(len(k)))
        assert abs(E - Eo) <= 1e-12 * max(1.0, len(k) * N * N / po.volume(c.basis)) + 1e-22
    assert len(out["success"]) == sum(1 for E in out["energies"] if E < 1e-14) >= S - 1


def test_potential_check_ratio_is_one():
    # Potential::Check (potential/Potential.cpp:5-19): one-sided difference, delta 1e-7
    cfg, k, v = stealthy(300, 3, 0.3, 13)
    pot = ShiftedCCPotential(3)
    for kk in k:
        pot.AddConstraint(kk, [1.0])
    for i in (0, 123):
        assert abs(pot.Check(cfg, i) - 1.0) < 1e-3


def test_mixed_precision_solves_end_in_fp64_ground_states():
    # FP32-phase descent, FP64 polish: every solve switches to FP64 and ends below Emin in FP64 (checked by the
    # oracle); the same job alone and in a mix gives the same bits
    specs = [stealthy(400, 2, 0.35, 41), stealthy(300, 3, 0.4, 42), stealthy(900, 2, 0.25, 43)]
    solver = BatchSolver(max_slots=8, precision="mixed", polish=1e-8)
    ids = [solver.submit(*sp, Emin=EMIN) for sp in specs]
    solver.run(timeout=300)
    for j, (cfg, k, v) in zip(ids, specs):
        r = solver.result(j)
        assert r["state"] == BatchSolver.DONE and r["energy"] < EMIN
        assert 0 < r["switch_eval"] <= r["device_evaluations"]
        Eo, _ = po.total_energy_forces(r["relative"], cfg.basis, k, v)
        assert Eo < 10 * EMIN
    alone = BatchSolver(max_slots=2, precision="mixed")
    ja = alone.submit(*specs[1], Emin=EMIN)
    alone.run(timeout=300)
    ra, rm = alone.result(ja), solver.result(ids[1])
    assert ra["energy"] == rm["energy"] and ra["switch_eval"] == rm["switch_eval"]
    assert np.array_equal(ra["relative"], rm["relative"])


def test_one_and_two_lanes_give_the_same_bits():
    # jobs spread over two streams (E+F of one beside L-BFGS of the other) or on one stream: same results
    specs = [stealthy(300, 2, 0.3, 51), stealthy(250, 3, 0.4, 52), stealthy(500, 2, 0.2, 53), stealthy(150, 3, 0.45, 54)]
    out = []
    for lanes in (1, 2):
        solver = BatchSolver(max_slots=8, lanes=lanes)
        ids = [solver.submit(*sp, Emin=EMIN) for sp in specs]
        solver.run(timeout=300)
        out.append([solver.result(j) for j in ids])
    for a, b in zip(*out):
        assert a["state"] == b["state"] == BatchSolver.DONE
        assert a["energy"] == b["energy"] and a["evaluations"] == b["evaluations"]
        assert np.array_equal(a["relative"], b["relative"])


def test_serving_stream_results_match_alone():
    # continuous batching: requests arrive while others are on the GPU and join the running batch; each result is
    # the same bits as the same request solved alone
    from modules_amd.serving import make_stream, serve
    reqs = make_stream(6, rate=200.0, seed=7, median_N=300, N_clip=(100, 700))
    reqs, wall = serve(reqs, Emin=EMIN, max_slots=4, timeout=300)
    assert all(r.result["state"] == BatchSolver.DONE and r.result["energy"] < EMIN for r in reqs)
    # some requests were on the GPU together (one arrived before another finished)
    assert any(a.submitted < b.done and b.submitted < a.done for a in reqs for b in reqs if a is not b)
    for r in reqs[:3]:
        cfg = RandomConfiguration(r.seed, r.N, r.d)
        k = GetKs(cfg, getK(r.chi, 1.0, r.d))
        alone = BatchSolver(max_slots=1)
        j = alone.submit(cfg, k, np.ones(len(k)), epsilon=r.epsilon, sigma=r.sigma, Emin=EMIN)
        alone.run(timeout=300)
        a = alone.result(j)
        assert a["energy"] == r.result["energy"] and a["evaluations"] == r.result["evaluations"]
        assert np.array_equal(a["relative"], r.result["relative"])
