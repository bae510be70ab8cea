"""Failure semantics of the batched solver: a job whose evaluation goes non-finite ends alone as failed (the
reference's NotANumberFound, potential/StructureOptimization.cpp:458-461), and a job past its wall-clock limit
returns its last evaluated configuration with NLopt's time-limit code (the per-job counterpart of the reference's
TimeLimit, :352-378).  Neighbouring jobs keep the bits they have alone."""
import numpy as np
import pytest

from modules_amd import GetKs, RandomConfiguration, getK
from modules_amd._lib import ERR_NAN
from modules_amd.optimize import BatchSolver
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

EMIN = 1e-18


def stealthy(N, d, chi, seed):
    cfg = RandomConfiguration(seed, N, d)
    k = GetKs(cfg, getK(chi, 1.0, d))
    return cfg, k, np.ones(len(k))


def solve_alone(spec):
    s = BatchSolver(max_slots=1)
    j = s.submit(*spec, Emin=EMIN)
    s.run(timeout=300)
    return s.result(j)


def test_overflowing_job_fails_and_its_neighbour_is_unaffected():
    good = stealthy(300, 2, 0.3, 1)
    cfg, k, _ = stealthy(250, 3, 0.35, 2)
    bad = (cfg, k, np.full(len(k), 1e308))  # 2 V / vol overflows: the force coefficients are not finite
    solver = BatchSolver(max_slots=4)
    jg = solver.submit(*good, Emin=EMIN)
    jb = solver.submit(*bad, Emin=EMIN)
    solver.run(timeout=300)
    rb = solver.result(jb)
    assert rb["state"] == BatchSolver.FAILED
    assert rb["code"] == ERR_NAN and "NotANumberFound" in rb["error"]
    assert rb["device_evaluations"] == 1  # the first evaluation already has a non-finite gradient
    rg = solver.result(jg)
    assert rg["state"] == BatchSolver.DONE and rg["energy"] < EMIN
    ra = solve_alone(good)
    assert ra["energy"] == rg["energy"] and ra["evaluations"] == rg["evaluations"]
    assert np.array_equal(ra["relative"], rg["relative"])
    # the failed job's slot is free again
    solver.release(jb)
    solver.release(jg)
    assert solver.stats()["bytes_in_use"] == 0


def test_time_limit_returns_the_last_evaluated_point():
    small = stealthy(300, 2, 0.3, 3)
    big = stealthy(20000, 2, 0.4, 4)
    solver = BatchSolver(max_slots=4)
    js = solver.submit(*small, Emin=EMIN)  # no limit
    solver.set_time_limit(0.05)
    jb = solver.submit(*big, Emin=EMIN)
    solver.run(timeout=300)
    rb = solver.result(jb)
    assert rb["state"] == BatchSolver.DONE and rb["status"] == 6  # NLOPT_MAXTIME_REACHED
    assert rb["energy"] > EMIN and 0 < rb["device_evaluations"] < 1400
    cfg, k, v = big
    # the returned energy is the oracle's energy of the returned configuration (scale-aware 1e-12)
    Eo, _ = po.total_energy_forces(rb["relative"], cfg.basis, k, v)
    scale = len(k) * cfg.NumParticle() ** 2 / cfg.PeriodicVolume()
    assert abs(Eo - rb["energy"]) <= 1e-12 * scale
    rs = solver.result(js)
    ra = solve_alone(small)
    assert rs["status"] != 6 and rs["energy"] == ra["energy"] and np.array_equal(rs["relative"], ra["relative"])


@pytest.mark.parametrize("N,d,chi", [(50000, 3, 0.49), (50000, 2, 0.49), (20000, 2, 0.4), (4000, 3, 0.45), (150, 2, 0.2)])
def test_admission_estimate_covers_the_resident_buffers(N, d, chi):
    # the HBM estimate of a job (same segment plan and chunk rule as the system's own preparation) is at least what
    # the resident job holds, up to the largest serving requests
    cfg, k, v = stealthy(N, d, chi, 9)
    solver = BatchSolver(max_slots=1)
    j = solver.submit(cfg, k, v, Emin=EMIN)
    solver.step(0)  # admission only
    est, held = solver.job_bytes(j)
    assert held > 0 and est >= held, (est, held)
    assert est <= 2 * held + (1 << 20)  # and not wildly above it
