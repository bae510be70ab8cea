"""GPU parity of molecular dynamics (ParticleMolecularDynamics, potential/MD_System.cpp) against the oracle's
restatement over its own forces (oracle/md.py): velocity Verlet, Nose-Hoover and Anderson trajectories over short
runs (the dynamics is chaotic, so runs stay short enough that rounding has not grown: 1e-9 relative), the random
draws consumed in the reference's order, and energy conservation over a longer NVE run."""
import numpy as np
import pytest

from modules_amd import GetKs, RandomConfiguration, RepulsiveCCPotential, ShiftedCCPotential, getK
from modules_amd.md import CollectiveCoordinateMD, ParticleMolecularDynamics
from modules_amd.potential import RandomGenerator
from oracle import md as omd
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def setup(N=200, d=2, chi=0.3, seed=3, softcore=True):
    cfg = RandomConfiguration(seed, N, d)
    k = GetKs(cfg, getK(chi, 1.0, d))
    if softcore:
        pot = RepulsiveCCPotential(d, 1.0, 0.5)
        for kk in k:
            pot.CCPotential.AddConstraint(kk, [1.0])
    else:
        pot = ShiftedCCPotential(d)
        for kk in k:
            pot.AddConstraint(kk, [1.0])
    e0, sigma = (1.0, 0.5) if softcore else (0.0, 0.0)

    def forces(s):
        return po.total_energy_forces(s, cfg.basis, k, np.ones(len(k)), 0.0, sigma, e0)[1]

    return cfg, k, pot, forces


def close(a, b, rel=1e-9):
    return np.max(np.abs(a - b)) <= rel * max(1.0, np.max(np.abs(b)))


def wrapped_close(s, t, tol=1e-9):
    d = s - t
    return np.max(np.abs(d - np.round(d))) <= tol


@pytest.mark.parametrize("d,softcore", [(2, True), (3, False)])
def test_velocity_verlet_matches_oracle(d, softcore):
    cfg, k, pot, forces = setup(N=150 if d == 2 else 120, d=d, softcore=softcore)
    md = ParticleMolecularDynamics(cfg, 0.01, 1.0)
    ref = omd.MD(cfg.basis, cfg.relative, 0.01, 1.0, forces)
    gen = RandomGenerator(7)
    md.SetRandomSpeed(0.5, gen)
    draws = RandomGenerator(7).RandomDoubles(10000)
    used = ref.set_random_speed(0.5, draws)
    assert gen._pos == used
    assert close(md.Velocity, ref.v, 1e-15)
    md.RemoveOverallMomentum()
    ref.remove_overall_momentum()
    assert np.allclose(md.Velocity.mean(axis=0), 0.0, atol=1e-14)
    md.Evolve(20, pot)
    ref.evolve(20)
    assert wrapped_close(md.Position.relative, ref.s)
    assert close(md.Velocity, ref.v) and close(md.Acceleration, ref.acc)
    assert md.Time == pytest.approx(ref.time, rel=1e-15)
    assert md.GetKineticEnergy() == pytest.approx(ref.kinetic(), rel=1e-9)
    # the potential is left bound to the final positions
    E, _ = po.total_energy_forces(ref.s, cfg.basis, k, np.ones(len(k)), 0.0, *((0.5, 1.0) if softcore else (0.0, 0.0)))
    assert pot.Energy() == pytest.approx(E, rel=1e-9, abs=1e-12)


def test_nose_hoover_matches_oracle():
    cfg, k, pot, forces = setup(N=120, d=2)
    md = ParticleMolecularDynamics(cfg, 0.01, 1.0)
    ref = omd.MD(cfg.basis, cfg.relative, 0.01, 1.0, forces)
    gen = RandomGenerator(11)
    md.SetRandomSpeed(0.3, gen)
    ref.set_random_speed(0.3, RandomGenerator(11).RandomDoubles(10000))
    md.InitNoseHoover(50.0)
    ref.Q = 50.0
    md.NoseHooverEvolve(15, pot, 0.2)
    ref.nose_hoover(15, 0.2)
    assert wrapped_close(md.Position.relative, ref.s)
    assert close(md.Velocity, ref.v)
    assert md.xi == pytest.approx(ref.xi, rel=1e-9, abs=1e-12)


def test_anderson_matches_oracle_and_consumes_the_reference_draws():
    cfg, k, pot, forces = setup(N=100, d=2)
    md = ParticleMolecularDynamics(cfg, 0.01, 1.0)
    ref = omd.MD(cfg.basis, cfg.relative, 0.01, 1.0, forces)
    gen = RandomGenerator(5)
    md.AndersonEvolve(25, pot, 0.4, 0.3, gen)
    used = ref.anderson(25, 0.4, 0.3, RandomGenerator(5).RandomDoubles(10000))
    assert gen._pos == used and used > 0
    assert wrapped_close(md.Position.relative, ref.s)
    assert close(md.Velocity, ref.v)


def test_nve_conserves_energy():
    cfg, k, pot, _ = setup(N=200, d=2, softcore=False)
    md = ParticleMolecularDynamics(cfg, 0.005, 1.0)
    md.SetRandomSpeed(0.05, RandomGenerator(3))
    E0 = md.GetKineticEnergy() + md.GetPotentialEnergy(pot)
    md.Evolve(400, pot)
    E1 = md.GetKineticEnergy() + md.GetPotentialEnergy(pot)
    print(f"NVE 400 steps: E {E0:.10g} -> {E1:.10g}", flush=True)
    assert abs(E1 - E0) <= 1e-3 * abs(E0)


def test_collective_coordinate_md_writes_samples(tmp_path):
    from modules_amd.formats import ConfigurationPack
    cfg, k, pot, _ = setup(N=60, d=2)
    out = CollectiveCoordinateMD(cfg, pot, RandomGenerator(1), 0.01, 0.1, str(tmp_path / "md"), 3, 20, 2)
    assert len(out) == 3 and all(np.isfinite(e) and ek > 0 for e, ek in out)
    assert ConfigurationPack(str(tmp_path / "md")).NumConfig() == 3


@pytest.mark.parametrize("d,N,sigma", [(2, 3000, 0.6), (3, 1500, 0.9), (3, 800, 1.6)])
def test_softcore_kernels_give_identical_bits(d, N, sigma):
    """The thread-per-particle and lane-group pair kernels sum every particle's pairs in the same tree."""
    from modules_amd._lib import lib
    cfg = RandomConfiguration(21 + d, N, d)
    pot = RepulsiveCCPotential(d, 1.0, sigma)
    results = []
    try:
        for mode in (1, 2):
            assert lib().ccgpu_set_softcore_kernel(mode) == 0
            pot.SetConfiguration(cfg)
            results.append(pot.EnergyAndAllForce())
    finally:
        lib().ccgpu_set_softcore_kernel(0)
    (E1, F1), (E2, F2) = results
    assert E1 == E2 and np.array_equal(F1, F2)
    Eo = po.softcore_energy(cfg.relative, cfg.basis, sigma, 1.0)
    Fo = po.softcore_forces(cfg.relative, cfg.basis, sigma, 1.0)
    assert E1 == pytest.approx(Eo, rel=1e-12) and np.allclose(F1, Fo, rtol=1e-11, atol=1e-12)


def test_mddump_restores_the_state_and_its_dynamics(tmp_path):
    # WriteBinary / ParticleMolecularDynamics(std::istream &) (potential/MD_System.cpp:34-62): a thermostatted run
    # written to the restart file and read back holds the same state, and both copies then evolve to the same bits
    cfg, k, pot_a, _ = setup(N=120, d=2)
    _, _, pot_b, _ = setup(N=120, d=2)
    md = ParticleMolecularDynamics(cfg, 0.01, np.linspace(0.8, 1.2, 120))
    md.SetRandomSpeed(0.2, RandomGenerator(3))
    md.InitNoseHoover(2.5)
    md.NoseHooverEvolve(15, pot_a, 0.2)
    path = tmp_path / "run.MDDump"
    md.WriteDump(path, 2, 41)
    stage, step, back = ParticleMolecularDynamics.ReadDump(path)
    assert (stage, step) == (2, 41)
    assert np.array_equal(back.Position.relative, md.Position.relative)
    assert np.array_equal(back.Velocity, md.Velocity) and np.array_equal(back.Acceleration, md.Acceleration)
    assert (back.Time, back.TimeStep, back.xi, back.MassQ) == (md.Time, md.TimeStep, md.xi, md.MassQ)
    assert md.xi != 0.0 and np.array_equal(back.Mass, md.Mass)
    md.NoseHooverEvolve(10, pot_a, 0.2)
    back.NoseHooverEvolve(10, pot_b, 0.2)
    assert np.array_equal(back.Position.relative, md.Position.relative)
    assert np.array_equal(back.Velocity, md.Velocity) and back.xi == md.xi


def test_collective_coordinate_md_continues_from_its_restart_file(tmp_path, monkeypatch):
    # CollectiveCoordinateMD with AllowRestore (apps/CCOptimization/main.cpp:1891-1911, 1936-1946, 2005-2011): a run
    # past its time limit stores stage 1 and the step about to run and says so in Prefix_continue.txt; the next
    # run continues from the step after it, keeps the pack, and stores stage 2 and the last sample at the end
    from modules_amd.formats import ConfigurationPack
    monkeypatch.chdir(tmp_path)
    cfg, k, pot, _ = setup(N=60, d=2)
    prefix = str(tmp_path / "md")
    out = CollectiveCoordinateMD(cfg, pot, RandomGenerator(1), 0.01, 0.1, prefix, 3, 20, 2, AllowRestore=True,
                                 TimeLimit=0.0)
    assert out == []
    assert open(prefix + "_continue.txt").read() == "Not Completed. Please run again\n"
    stage, step, _ = ParticleMolecularDynamics.ReadDump(prefix + ".MDDump")
    assert (stage, step) == (1, 0)
    out = CollectiveCoordinateMD(cfg, pot, RandomGenerator(1), 0.01, 0.1, prefix, 3, 20, 2, AllowRestore=True)
    assert len(out) == 3 and all(np.isfinite(e) and ek > 0 for e, ek in out)
    assert ConfigurationPack(prefix).NumConfig() == 3
    stage, step, system = ParticleMolecularDynamics.ReadDump(prefix + ".MDDump")
    assert (stage, step) == (2, 2)
    assert np.array_equal(system.Position.relative, ConfigurationPack(prefix).GetConfig(2).relative)
    # a run that finds the finished file takes no more samples and keeps the pack
    assert CollectiveCoordinateMD(cfg, pot, RandomGenerator(1), 0.01, 0.1, prefix, 3, 20, 2, AllowRestore=True) == []
    assert ConfigurationPack(prefix).NumConfig() == 3
