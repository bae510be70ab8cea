"""GPU parity of the multi-species potential (MultiStealthyPotential) against the oracle.

Cases follow the reference's fused-contact test table (apps/CCOptimization/tests/test_fused_contact.cpp:387-418):
dimensions 1-3, square and skewed cells, 1-10 species, unequal radii, coincident particles, tiny cells where a
pair touches through several periodic images, contact off.  Tolerance: FP64 within 1e-12 relative to the energy
scale (species CC sums plus the contact count) and 1e-11 relative to the largest force component, as the
single-species soft-core test.  Bitwise checks follow the reference's own: EnergyAndAllForce equals Energy then
AllForce, repeated evaluations and a fresh potential repeat the bits, reassignment equals direct assignment.
"""
import numpy as np
import pytest

from modules_amd import Configuration, MultiStealthyPotential, SpherePacking
from modules_amd._lib import CCGPUError, ERR_STATE
from multi_cases import cell, radii_of, species_setup
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

# (dim, skew, sizes, chi, e0, phi, unequal, coincident)
CASES = [
    (1, False, (400,), 0.3, 1.0, 0.2, False, 0),
    (1, False, (200, 150, 100), 0.5, 1.0, 0.4, True, 3),
    (2, True, (300, 500), 0.25, 1.0, 0.3, True, 0),
    (2, False, (200,) * 5, 0.4, 0.0, 0.2, False, 0),
    (3, False, (150,) * 5, 0.2, 1.0, 0.3, True, 4),
    (3, True, (90, 100, 110, 120, 130, 140, 150, 160, 170, 180), 0.2, 1.0, 0.2, False, 0),
    (2, True, (150,) * 6, 0.2, 1.0, 0.2, True, 2),
    (3, True, (500, 400), 1.0, 1.0, 0.2, False, 0),
    # tiny cells: contact through several images of a neighbour and of the particle itself
    (1, False, (3, 2), 0.5, 1.0, 3.0, True, 1),
    (2, True, (4, 3), 0.5, 1.0, 3.0, True, 1),
    (3, False, (5, 4), 0.5, 1.0, 3.0, True, 1),
    (3, True, (5, 4), 0.5, 0.0, 3.0, True, 1),
] + [(2, S % 2 == 0, (40 + 3 * S,) * S, 0.4, 1.0 if S % 3 else 0.0, 0.3, True, S % 4) for S in (1, 2, 3, 4, 7, 10)]


def build(case, seed_shift=0):
    d, skew, sizes, chi, e0, phi, unequal, coincident = case
    N = sum(sizes)
    a = cell(N, d, skew)
    s = po.make_config(31 + N + d + seed_shift, N, d)
    for q in range(coincident):
        s[N - 1 - q] = s[q]
    radii = radii_of(N, d, phi, unequal)
    idx, ks, vs = species_setup(a, sizes, chi)
    return Configuration(a, s), radii, idx, ks, vs, e0


def make_potential(cfg, radii, idx, ks, vs, e0):
    d = cfg.GetDimension()
    p = MultiStealthyPotential(d, e0)
    p.SetConfiguration(cfg)
    p.SetNumSpecies(len(idx))
    p.SetSecureRadii(radii)
    label = np.empty(cfg.NumParticle(), dtype=int)
    for q, lst in enumerate(idx):
        label[lst] = q
    for i, q in enumerate(label):  # particle order, as the reference test assigns
        p.SetSpeciesOfParticle(i, int(q))
    for q in range(len(idx)):
        for k, v in zip(ks[q], vs[q]):
            p.AddConstraint(q, k, [v])
    p.SetConfiguration(cfg)
    return p


def oracle(cfg, radii, idx, ks, vs, e0, u0=1.0, repulsion=None):
    rep = abs(e0) > 1e-10 if repulsion is None else repulsion
    return po.multi_energy_forces(cfg.relative, cfg.basis, idx, ks, vs, radii, 2.01 * float(np.max(radii)), u0,
                                  e0, rep)


def energy_scale(cfg, idx, vs, e0, u0=1.0):
    N = cfg.NumParticle()
    vol = po.volume(cfg.basis)
    cc = sum(len(lst) / N * np.sum(np.abs(v)) * len(lst) ** 2 / vol for lst, v in zip(idx, vs))
    return abs(u0) * cc + abs(e0) * N + 1.0


def check(E, F, Eo, Fo, scale):
    assert abs(E - Eo) <= 1e-12 * scale, (E, Eo, scale)
    assert np.max(np.abs(F - Fo)) <= 1e-11 * max(1.0, np.max(np.abs(Fo)))


@pytest.mark.parametrize("case", CASES, ids=[f"d{c[0]}{'s' if c[1] else ''}-S{len(c[2])}-N{sum(c[2])}-e{c[4]}-"
                                             f"phi{c[5]}-c{c[7]}" for c in CASES])
def test_cases_match_oracle(case):
    cfg, radii, idx, ks, vs, e0 = build(case)
    print(f"multi case {case}", flush=True)
    p = make_potential(cfg, radii, idx, ks, vs, e0)
    assert [sp.idx for sp in p.species_constraints] == [list(map(int, lst)) for lst in idx]
    E, F = p.EnergyAndAllForce()
    Eo, Fo, ES = oracle(cfg, radii, idx, ks, vs, e0)
    check(E, F, Eo, Fo, energy_scale(cfg, idx, vs, e0))
    vol = po.volume(cfg.basis)
    for q, sp in enumerate(p.species_constraints):
        n_s = len(idx[q])
        assert abs(sp.CCPotential.Energy() - ES[q]) <= 1e-12 * (np.sum(vs[q]) * n_s * n_s / vol + 1.0)
        p.GetRho(q)
        re, im = po.rho(cfg.relative[idx[q]], cfg.basis, ks[q])
        assert np.max(np.abs(sp.CCPotential.RhoReal - re), initial=0.0) <= 1e-12 * max(n_s, 1)
        assert np.max(np.abs(sp.CCPotential.RhoImag - im), initial=0.0) <= 1e-12 * max(n_s, 1)


def test_fused_equals_separate_and_bits_repeat():
    case = CASES[2]
    cfg, radii, idx, ks, vs, e0 = build(case)
    a = make_potential(cfg, radii, idx, ks, vs, e0)
    E1 = a.Energy()
    F1 = a.AllForce()
    b = make_potential(cfg, radii, idx, ks, vs, e0)
    E2, F2 = b.EnergyAndAllForce()
    assert E1 == E2 and np.array_equal(F1, F2)
    c = make_potential(cfg, radii, idx, ks, vs, e0)
    F3 = c.AllForce()  # cold AllForce (MD order)
    assert np.array_equal(F3, F1) and c.Energy() == E1
    for i in (0, 17, cfg.NumParticle() - 1):
        assert np.array_equal(a.Force(i), F1[i])
    a.SetConfiguration(cfg)
    E4, F4 = a.EnergyAndAllForce()
    assert E4 == E1 and np.array_equal(F4, F1)


def test_reassignment_equals_direct_assignment():
    case = (2, True, (60, 80, 40), 0.3, 1.0, 0.3, True, 0)
    cfg, radii, idx, ks, vs, e0 = build(case)
    p = make_potential(cfg, radii, idx, ks, vs, e0)
    p.EnergyAndAllForce()
    # move particles between species, as SetSpeciesOfParticle does (removal keeps the others' order)
    moves = [(idx[0][3], 2), (idx[2][0], 1), (idx[1][5], 0), (idx[0][10], 2)]
    lists = [list(map(int, lst)) for lst in idx]
    for i, q in moves:
        p.SetSpeciesOfParticle(int(i), q)
        for lst in lists:
            if int(i) in lst:
                lst.remove(int(i))
        lists[q].append(int(i))
    assert [sp.idx for sp in p.species_constraints] == lists
    E, F = p.EnergyAndAllForce()
    # a fresh potential given the final lists directly (same species orders) repeats the bits
    q = MultiStealthyPotential(2, e0)
    q.SetConfiguration(cfg)
    q.SetNumSpecies(3)
    q.SetSecureRadii(radii)
    for s, lst in enumerate(lists):
        for i in lst:
            q.SetSpeciesOfParticle(i, s)
        for k, v in zip(ks[s], vs[s]):
            q.AddConstraint(s, k, [v])
    Eq, Fq = q.EnergyAndAllForce()
    assert E == Eq and np.array_equal(F, Fq)
    Eo, Fo, _ = oracle(cfg, radii, lists, ks, vs, e0)
    check(E, F, Eo, Fo, energy_scale(cfg, lists, vs, e0))


def test_scales_radii_and_packing():
    case = (3, False, (70, 50), 0.3, 1.0, 0.25, True, 1)
    cfg, radii, idx, ks, vs, e0 = build(case)
    p = make_potential(cfg, radii, idx, ks, vs, e0)
    p.u0, p.e0 = 0.6, 2.5
    E, F = p.EnergyAndAllForce()
    Eo, Fo, _ = oracle(cfg, radii, idx, ks, vs, 2.5, u0=0.6)
    check(E, F, Eo, Fo, energy_scale(cfg, idx, vs, 2.5, 0.6))
    # radii increased beyond the original cutoff
    big = radii * 1.6
    p.SetSecureRadii(big)
    assert p.GetRc() == pytest.approx(2.01 * big.max(), rel=1e-15)
    E, F = p.EnergyAndAllForce()
    Eo, Fo, _ = oracle(cfg, big, idx, ks, vs, 2.5, u0=0.6)
    check(E, F, Eo, Fo, energy_scale(cfg, idx, vs, 2.5, 0.6))
    # SetPacking carries the radii
    pck = SpherePacking(cfg.basis, cfg.relative, big)
    p.SetPacking(pck)
    E2, F2 = p.EnergyAndAllForce()
    assert E2 == E and np.array_equal(F2, F)
    with pytest.raises(ValueError):
        p.SetSecureRadii(np.where(np.arange(len(big)) == 3, -1.0, big))


def test_contact_strength_fixed_at_creation():
    # WithRepulsion is decided by the constructor's epsilon; setting e0 later only scales (MultiStealthyPotential.h)
    case = (2, False, (50, 50), 0.3, 0.0, 0.4, True, 0)
    cfg, radii, idx, ks, vs, _ = build(case)
    p = make_potential(cfg, radii, idx, ks, vs, 0.0)
    p.e0 = 1.0
    E, F = p.EnergyAndAllForce()
    Eo, Fo, _ = oracle(cfg, radii, idx, ks, vs, 1.0, repulsion=False)
    check(E, F, Eo, Fo, energy_scale(cfg, idx, vs, 1.0))


def test_unassigned_particle_is_an_error():
    cfg, radii, idx, ks, vs, e0 = build((2, False, (30, 30), 0.3, 1.0, 0.2, False, 0))
    p = MultiStealthyPotential(2, 1.0)
    p.SetConfiguration(cfg)
    p.SetNumSpecies(2)
    for i in idx[0]:
        p.SetSpeciesOfParticle(int(i), 0)
    with pytest.raises(CCGPUError) as e:
        p.Energy()
    assert e.value.code == ERR_STATE
    assert list(p.SpeciesOf()[idx[1]]) == [-1] * len(idx[1])


def test_relax_two_species_lowers_energy_to_a_ground_state():
    case = (2, False, (60, 60), 0.2, 1.0, 0.2, False, 0)
    cfg, radii, idx, ks, vs, e0 = build(case)
    p = make_potential(cfg, radii, idx, ks, vs, e0)
    E0 = p.Energy()
    E, evals, status = p.Relax(maxstep=10000, Emin=1e-20)
    print(f"multi relax: E0={E0:.6g} E={E:.6g} evaluations={evals} status={status}", flush=True)
    assert evals > 0
    assert E < 1e-10 * E0
    # the relaxed configuration stays bound: a fresh evaluation gives the same energy
    assert p.Energy() == E
    Eo, Fo, _ = oracle(p._config, radii, idx, ks, vs, e0)
    assert abs(E - Eo) <= 1e-12 * energy_scale(p._config, idx, vs, e0)


def mixture_job(seed, sizes=(60, 60), chi=0.1, d=2, phi=0.2):
    cfg, radii, idx, ks, vs, e0 = build((d, False, sizes, chi, 1.0, phi, False, 0), seed_shift=seed)
    return cfg, make_potential(cfg, radii, idx, ks, vs, e0), (radii, idx, ks, vs, e0)


def test_solver_mixture_jobs_reach_ground_states_beside_single_species_jobs():
    from modules_amd import GetKs, RandomConfiguration, getK
    from modules_amd.optimize import BatchSolver
    jobs = [mixture_job(s) for s in range(3)] + [mixture_job(7, sizes=(40, 50, 30), d=3)]
    solver = BatchSolver(max_slots=8, check_interval=8)
    ids = [solver.submit_multi(cfg, p, Emin=1e-20) for cfg, p, _ in jobs]
    single = RandomConfiguration(3, 200, 2)
    ks = GetKs(single, getK(0.3, 1.0, 2))
    sid = solver.submit(single, ks, np.ones(len(ks)), Emin=1e-20)
    solver.run(timeout=300)
    for (cfg, p, (radii, idx, kk, vv, e0)), j in zip(jobs, ids):
        r = solver.result(j)
        assert r["state"] == BatchSolver.DONE
        out = Configuration(cfg.basis, r["relative"])
        p.SetConfiguration(out)
        E = p.Energy()
        print(f"solver mixture job: E={r['energy']:.4g} evaluations={r['evaluations']} status={r['status']}",
              flush=True)
        assert abs(E - r["energy"]) <= 1e-12 * energy_scale(out, idx, vv, e0)
        assert r["energy"] < 1e-14
        Eo, _, _ = oracle(out, radii, idx, kk, vv, e0)
        assert abs(Eo - r["energy"]) <= 1e-12 * energy_scale(out, idx, vv, e0)
    assert solver.result(sid)["energy"] < 1e-18


def test_solver_mixture_job_same_bits_alone_and_in_a_mix():
    from modules_amd.optimize import BatchSolver
    cfg, p, _ = mixture_job(11, sizes=(50, 40, 30))
    alone = BatchSolver(max_slots=2, check_interval=8, lanes=1)
    a = alone.submit_multi(cfg, p, MaxStep=300)
    alone.run(timeout=300)
    ra = alone.result(a)
    mix = BatchSolver(max_slots=6, check_interval=8, lanes=2)
    others = [mixture_job(s)[0:2] for s in (12, 13)]
    for c, q in others[:1]:
        mix.submit_multi(c, q, MaxStep=300)
    b = mix.submit_multi(cfg, p, MaxStep=300)
    for c, q in others[1:]:
        mix.submit_multi(c, q, MaxStep=300)
    mix.run(timeout=300)
    rb = mix.result(b)
    assert ra["energy"] == rb["energy"] and ra["evaluations"] == rb["evaluations"]
    assert np.array_equal(ra["relative"], rb["relative"])


def test_multirun_with_a_mixture():
    from modules_amd.optimize import CollectiveCoordinateMultiRun
    from modules_amd.potential import RandomGenerator
    cfg, p, _ = mixture_job(21)
    gen = RandomGenerator(5)
    out = CollectiveCoordinateMultiRun(cfg, p, gen, SampleNumber=3, E_max=1e-14, max_num=10000)
    print("multirun mixture energies:", out["energies"], flush=True)
    assert len(out["relaxed"]) == 3
    assert len(out["success"]) == sum(e < 1e-14 for e in out["energies"]) >= 2


@pytest.mark.parametrize("dim", [2, 3])
def test_export_configuration_follows_the_reference_export_test(dim, tmp_path):
    # apps/CCOptimization/tests/test_export.cpp: errors before a configuration and before every particle has a
    # species; after the reassignments the text and MSPB bytes carry the stored configuration, labels and radii
    from modules_amd._lib import CCGPUError
    from modules_amd.formats import ExportText, MSPBSnapshot
    a = np.zeros((dim, dim))
    for d in range(dim):
        a[d, d] = 4.0 + d
    a[1, 0] = 1.0 / 7.0
    s = np.array([[(i + d + 1) / 11.0 for d in range(dim)] for i in range(3)])
    p = MultiStealthyPotential(dim, 1.0)
    with pytest.raises(CCGPUError):
        p.ExportConfiguration(None)
    p.SetConfiguration(Configuration(a, s))
    p.SetNumSpecies(2)
    with pytest.raises(CCGPUError):
        p.ExportConfiguration(None, "binary")
    for i, q in ((0, 0), (0, 1), (1, 1), (2, 0)):
        p.SetSpeciesOfParticle(i, q)
    p.SetSecureRadii([0.125, 0.25, 0.375])
    text = p.ExportConfiguration(str(tmp_path / "c.txt"))
    assert text == ExportText(a, s, [1, 1, 0], [0.125, 0.25, 0.375])
    assert open(tmp_path / "c.txt").read() == text
    data = p.ExportConfiguration(None, "binary")
    assert data == MSPBSnapshot(a, s, [1, 1, 0], [0.125, 0.25, 0.375], 2).Write()
    with pytest.raises(ValueError):
        p.ExportConfiguration(None, "xml")


def test_multirun_writes_the_reference_packs(tmp_path):
    from modules_amd.formats import ConfigurationPack
    from modules_amd.optimize import CollectiveCoordinateMultiRun
    from modules_amd.potential import RandomGenerator
    cfg, p, _ = mixture_job(31)
    prefix = str(tmp_path / "run")
    seen = []
    out = CollectiveCoordinateMultiRun(cfg, p, RandomGenerator(9), Prefix=prefix, SampleNumber=2, E_max=1e-14,
                                       onAccepted=lambda c, i: seen.append(i))
    relaxed, success, init = (ConfigurationPack(prefix), ConfigurationPack(prefix + "_Success"),
                              ConfigurationPack(prefix + "_InitConfig"))
    assert relaxed.NumConfig() == init.NumConfig() == 2
    assert success.NumConfig() == len(out["success"]) and seen == list(range(len(out["success"])))
    for q in range(2):
        assert np.array_equal(relaxed.GetConfig(q).relative, out["relaxed"][q].relative)
        assert np.array_equal(init.GetConfig(q).relative, out["init"][q].relative)
