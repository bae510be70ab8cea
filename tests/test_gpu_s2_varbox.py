"""GPU parity of the target-S(k) potential (ShiftedCCPotential_S2) and of the variable-cell potentials
(CCPotential_varBox.h, CCPotential_Kernels.h) against the oracle.

Bar: FP64 within 1e-12 relative on scale-aware magnitudes, as for the plain form: rho ~ N; S2 energy
~ sum|V| (N + max|S0|)^2; S2 force ~ 4 sum|V| (N + max|S0|) kmax (the bounds of |S - S0| and |rho| |k| / N times
the weights).  The NUFFT path keeps its own bar (1e-9 relative at eps 1e-12).  Variable-cell k-point sets and
coefficients are compared exactly; the cell derivative (a host contraction of the GPU's rho, in a different
summation order from the reference's) to 1e-10 relative of its largest component.
"""
import numpy as np
import pytest

from modules_amd import Configuration, GetKs, RandomConfiguration, RepulsiveCCPotential, ShiftedCCPotential_S2, getK
from modules_amd import varbox as VB
from modules_amd._lib import CCGPUError
from modules_amd.optimize import BatchSolver, RelaxStructure_NLOPT
from oracle import pyoracle as po
from oracle import varbox as ovb

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def s2_scales(k, v, s0, N):
    sumV = np.sum(np.abs(v))
    kmax = np.max(np.linalg.norm(k, axis=1)) if len(k) else 0.0
    top = N + (np.max(np.abs(s0)) if len(s0) else 0.0)
    return max(N, 1.0), sumV * top * top, 4.0 * sumV * top * kmax


def s2_pot(d, k, v, s0, shift=0.0):
    pot = ShiftedCCPotential_S2(d)
    for kk, vv, ss in zip(k, v, s0):
        pot.AddConstraint(kk, [vv, ss])
    pot.Shift = shift
    return pot


def s2_compare(cfg, k, v, s0, shift=0.0, backend="direct", rtol=RTOL):
    d, N = cfg.GetDimension(), cfg.NumParticle()
    pot = s2_pot(d, k, v, s0, shift)
    pot.SetRhoBackend(backend)
    pot.SetConfiguration(cfg)
    pot.GetRho()
    E, F = pot.EnergyAndAllForce()
    re, im = po.rho(cfg.relative, cfg.basis, k)
    Eo = po.s2_energy(v, s0, re, im, N, shift)
    Fo = po.s2_forces(cfg.relative, cfg.basis, k, v, s0, re, im)
    sr, se, sf = s2_scales(k, v, s0, N)
    err = (max(np.max(np.abs(pot.RhoReal - re)), np.max(np.abs(pot.RhoImag - im))) / sr, abs(E - Eo) / se,
           np.max(np.abs(F - Fo)) / sf)
    print(f"S2 {backend} d={d} N={N} M={len(k)}: rho {err[0]:.2e} E {err[1]:.2e} F {err[2]:.2e}", flush=True)
    assert err[0] <= rtol and err[1] <= rtol and err[2] <= rtol
    return pot, E, F


@pytest.mark.parametrize("N,d,chi,seed", [(300, 2, 0.3, 0), (1000, 3, 0.45, 1), (200, 1, 0.4, 2)])
def test_s2_lattice_path_matches_oracle(N, d, chi, seed):
    cfg = RandomConfiguration(seed, N, d)
    k = GetKs(cfg, getK(chi, 1.0, d))
    rng = np.random.default_rng(seed)
    pot, _, _ = s2_compare(cfg, k, 0.5 + rng.random(len(k)), rng.random(len(k)) * 2.0, shift=0.25)
    assert pot.UsesLatticePath()


def test_s2_general_path_matches_oracle():
    cfg = RandomConfiguration(5, 400, 2)
    rng = np.random.default_rng(5)
    k = rng.normal(size=(60, 2)) * 1.7  # not on the reciprocal lattice
    pot, _, _ = s2_compare(cfg, k, 0.5 + rng.random(60), rng.random(60))
    assert not pot.UsesLatticePath()


def test_s2_nufft_path_matches_oracle():
    cfg = RandomConfiguration(6, 4000, 2)
    k = GetKs(cfg, getK(0.4, 1.0, 2))
    rng = np.random.default_rng(6)
    pot, _, _ = s2_compare(cfg, k, np.ones(len(k)), rng.random(len(k)), backend="nufft_always", rtol=1e-9)
    assert pot.GetLastRhoBackendUsed() == "nufft"


def test_s2_equiluminous_soft_core_matches_oracle():
    # RepulsiveCCPotential with S0 != 0 (main.cpp:302-318): every constraint {1, S0}, plus the soft core
    cfg = RandomConfiguration(8, 500, 2)
    k = GetKs(cfg, getK(0.3, 1.0, 2))
    pot = RepulsiveCCPotential(2, 1.0, 0.6, S0=0.5)
    for kk in k:
        pot.CCPotential.AddConstraint(kk, [1.0, 0.5])
    pot.SetConfiguration(cfg)
    E, F = pot.EnergyAndAllForce()
    re, im = po.rho(cfg.relative, cfg.basis, k)
    v, s0 = np.ones(len(k)), np.full(len(k), 0.5)
    Eo = po.s2_energy(v, s0, re, im, 500) + po.softcore_energy(cfg.relative, cfg.basis, 0.6, 1.0)
    Fo = po.s2_forces(cfg.relative, cfg.basis, k, v, s0, re, im) + \
        po.softcore_forces(cfg.relative, cfg.basis, 0.6, 1.0)
    _, se, sf = s2_scales(k, v, s0, 500)
    assert abs(E - Eo) <= RTOL * se and np.max(np.abs(F - Fo)) <= RTOL * sf


def test_s2_constraints_need_a_target():
    pot = ShiftedCCPotential_S2(2)
    with pytest.raises(ValueError):
        pot.AddConstraint([1.0, 0.0], [1.0])
    from modules_amd._lib import lib
    import ctypes as C
    k = np.array([1.0, 0.0])
    assert lib().ccgpu_potential_add_constraint(pot._h, k.ctypes.data_as(C.POINTER(C.c_double)), 1.0) == -1


def test_s2_relax_and_solver_jobs():
    d, N = 2, 120
    cfg = RandomConfiguration(9, N, d)
    k = GetKs(cfg, getK(0.25, 1.0, d))
    v = np.ones(len(k))
    s0 = np.full(len(k), 0.5)
    pot = s2_pot(d, k, v, s0)
    c1 = cfg.copy()
    pot.SetConfiguration(c1)
    E0 = pot.Energy()
    E1, ev, _ = RelaxStructure_NLOPT(c1, pot, 0.0, 0, 0.0, 2000)
    re, im = po.rho(c1.relative, c1.basis, k)
    assert E1 == pytest.approx(po.s2_energy(v, s0, re, im, N), rel=1e-10, abs=1e-14)
    assert E1 < 0.01 * E0 and ev > 0
    # the same request through the batched solver, beside a plain job
    sol = BatchSolver(max_slots=4)
    j1 = sol.submit_s2(cfg, k, v, s0, MaxStep=2000)
    j2 = sol.submit(cfg, k, v, MaxStep=2000)
    sol.run()
    r1, r2 = sol.result(j1), sol.result(j2)
    s1 = np.asarray(r1["relative"]).reshape(N, d)
    re, im = po.rho(s1, cfg.basis, k)
    assert r1["energy"] == pytest.approx(po.s2_energy(v, s0, re, im, N), rel=1e-10, abs=1e-14)
    assert r1["energy"] < 0.01 * E0
    s2 = np.asarray(r2["relative"]).reshape(N, d)
    re, im = po.rho(s2, cfg.basis, k)
    assert r2["energy"] == pytest.approx(po.energy(v, re, im, po.volume(cfg.basis)), rel=1e-10, abs=1e-14)


# ---------------------------------------------------------------------------------------------------------------
# variable-cell potentials
# ---------------------------------------------------------------------------------------------------------------
SPECS = [
    ("stealthy", "K2", None, 1.0),
    ("stealthy", "KM3", None, 1.0),
    ("S2", "K2", "K2", 1.0),
    ("S2", "KM3", "Kn", 1.5),
    ("S2", "KM2", "KMLog", 1.0),
    ("S2_noV0", "K2", "K2", 1.0),
    ("S2_noV0", "KM3", "Klc", 2.0),
]


def cells(d):
    if d == 2:
        return np.array([[9.0, 0.0], [1.3, 8.2]]), np.array([[9.4, 0.2], [1.1, 8.0]])
    return (np.array([[6.5, 0.0, 0.0], [0.7, 6.1, 0.0], [-0.4, 0.9, 6.8]]),
            np.array([[6.6, 0.1, 0.0], [0.6, 6.2, -0.2], [-0.3, 0.8, 6.7]]))


@pytest.mark.parametrize("form,vk,s0,n", SPECS)
@pytest.mark.parametrize("d", [2, 3])
def test_varbox_matches_oracle(form, vk, s0, n, d):
    rng = np.random.default_rng(20 + d)
    N = 40
    s = rng.random((N, d))
    pot = VB.VarBoxPotential(d, form, vk, s0, 1.0, n)
    spec = ovb.Spec(form, vk, s0, 1.0, n)
    for a in cells(d):  # the second cell regenerates the constraints and the Shift
        cfg = Configuration(a, s)
        pot.SetConfiguration(cfg)
        Eo, re, im, k, nn, v, t = ovb.energy(spec, a, s)
        assert len(pot.constraints) == len(k) > 0
        assert np.array_equal(pot.kCoeffs, nn)
        kg = np.array([c.k for c in pot.constraints])
        assert np.allclose(kg, k, rtol=0, atol=1e-14)
        assert np.allclose([c.V for c in pot.constraints], v, rtol=1e-13, atol=1e-15)
        if t is not None:
            assert np.allclose(pot.S0, t, rtol=1e-13, atol=1e-15)
        E, F = pot.EnergyAndAllForce()
        scale = np.sum(np.abs(v)) * (N + (np.max(np.abs(t)) if t is not None else 0.0)) ** 2 + abs(
            ovb.shift(form, N, po.volume(a), v, t))
        assert abs(E - Eo) <= RTOL * scale, (E, Eo)
        kmax = np.max(np.linalg.norm(k, axis=1))
        if form == "stealthy":
            Fo = po.forces(s, a, k, v, re, im)
            fscale = 4.0 * np.sum(np.abs(v)) * N * kmax / po.volume(a)
        else:
            Fo = po.s2_forces(s, a, k, v, t, re, im)
            fscale = s2_scales(k, v, t, N)[2]
        assert np.max(np.abs(F - Fo)) <= RTOL * fscale
        for P in (0.0, 0.37):
            g = pot.EnergyDerivativeToBasisVectors(P)
            go = ovb.derivative(spec, a, s, P)
            assert np.max(np.abs(g - go)) <= 1e-10 * max(1.0, np.max(np.abs(go))), (g, go)
    # the analytic derivative against central differences of the GPU's E + P V (Potential::Check2's test,
    # potential/Potential.cpp:21-40, two-sided)
    a = cells(d)[0]
    pot.SetConfiguration(Configuration(a, s))
    g = pot.EnergyDerivativeToBasisVectors(0.2)
    h = 1e-6
    for idx in range(d * d):
        Hs = []
        for sgn in (1, -1):
            b = a.copy()
            b[idx // d, idx % d] += sgn * h
            pot.SetConfiguration(Configuration(b, s))
            Hs.append(pot.Energy() + 0.2 * po.volume(b))
        fd = (Hs[0] - Hs[1]) / (2 * h)
        assert abs(fd - g[idx]) <= 2e-6 * max(1.0, np.max(np.abs(g))), (idx, fd, g[idx])


def test_varbox_rejects_what_the_reference_cannot_do():
    pot = VB.K2Potential(2)
    with pytest.raises(RuntimeError):
        pot.AddConstraint([1.0, 0.0], [1.0])
    with pytest.raises(CCGPUError):
        VB.VarBoxPotential(3, "stealthy", "Lemniscate", None, 1.0)
    with pytest.raises(CCGPUError):
        VB.VarBoxPotential(2, "S2", "K2", None, 1.0)
    plain = ShiftedCCPotential_S2(2)
    plain.AddConstraint([1.0, 0.0], [1.0, 0.5])
    plain.SetConfiguration(RandomConfiguration(1, 10, 2))
    from modules_amd._lib import lib
    g = np.zeros(4)
    import ctypes as C
    assert lib().ccgpu_potential_energy_derivative_to_basis(plain._h, 0.0, g.ctypes.data_as(
        C.POINTER(C.c_double))) == -2


def test_lemniscate_weights_on_the_gpu():
    pot = VB.LemniscateVkPotential(1.2)
    a = np.array([[11.0, 0.0], [0.4, 10.0]])
    s = np.random.default_rng(3).random((30, 2))
    pot.SetConfiguration(Configuration(a, s))
    spec = ovb.Spec("stealthy", "Lemniscate", None, 1.2)
    Eo, re, im, k, nn, v, t = ovb.energy(spec, a, s)
    assert np.array_equal(pot.kCoeffs, nn)
    assert pot.Energy() == pytest.approx(Eo, rel=1e-12)
    assert np.allclose(pot.EnergyDerivativeToBasisVectors(0.1), ovb.derivative(spec, a, s, 0.1), rtol=1e-10,
                       atol=1e-12)


RELAX_CASES = [  # (form, vk, s0, d, N, seed, cell, pressure, switch, maxstep)
    ("stealthy", "K2", None, 2, 24, 30, [[7.0, 0.0], [0.8, 6.5]], 0.5, 1, 400),
    ("stealthy", "K2", None, 2, 24, 30, [[7.0, 0.0], [0.8, 6.5]], 0.5, 2, 400),
    ("stealthy", "K2", None, 2, 30, 31, [[16.0, 0.0], [1.5, 15.0]], 0.02, 1, 300),
    ("S2", "K2", "K2", 2, 30, 31, [[16.0, 0.0], [1.5, 15.0]], 0.0, 2, 300),
    ("stealthy", "K2", None, 3, 20, 32, [[5.0, 0, 0], [0.4, 5.2, 0], [0.2, -0.3, 4.8]], 0.3, 2, 200),
]


@pytest.mark.parametrize("form,vk,s0,d,N,seed,cell,P,switch,maxstep", RELAX_CASES)
def test_variable_cell_relax_matches_oracle(form, vk, s0, d, N, seed, cell, P, switch, maxstep):
    """RelaxStructure_NLOPT with Switch 1 / 2 against the oracle's restatement driven by the reference's own
    LuksanPLIS: same trajectory (evaluation count), enthalpy, cell and particles."""
    s = np.random.default_rng(seed).random((N, d))
    a = np.array(cell, dtype=np.float64)
    cfg = Configuration(a, s)
    pot = VB.VarBoxPotential(d, form, vk, s0, 1.0)
    spec = ovb.Spec(form, vk, s0, 1.0)
    pot.SetConfiguration(cfg)
    H0 = pot.Energy() + P * cfg.PeriodicVolume()
    H, ev, st = RelaxStructure_NLOPT(cfg, pot, P, switch, 0.0, maxstep)
    Ho, ao, so, evo = ovb.relax(spec, a, s, P, switch, maxstep)
    print(f"{form} d={d} switch {switch}: H0 {H0:.6g} -> {H:.10g} (oracle {Ho:.10g}), {ev} / {evo} evaluations",
          flush=True)
    assert H <= H0
    assert ev == evo
    assert H == pytest.approx(Ho, rel=1e-9, abs=1e-12)
    assert np.allclose(cfg.basis, ao, rtol=0, atol=1e-7 * np.max(np.abs(ao)))
    ds = cfg.relative - so
    assert np.max(np.abs(ds - np.round(ds))) < 1e-7
    # the relaxed configuration is left bound, its constraints regenerated for the final cell
    assert np.array_equal(pot.kCoeffs, ovb.kpoints(ao, 1.0)[1])
    if switch == 1:
        assert np.array_equal(cfg.relative, s)


# ---------------------------------------------------------------------------------------------------------------
# weighted rho and the log form
# ---------------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("N,d,chi,seed,general", [(400, 2, 0.3, 3, False), (1500, 3, 0.4, 4, False),
                                                  (300, 2, 0.3, 5, True)])
def test_weighted_rho_matches_oracle(N, d, chi, seed, general):
    from modules_amd import ShiftedCCPotential, ShiftedCCPotential_withWeight
    cfg = RandomConfiguration(seed, N, d)
    rng = np.random.default_rng(seed)
    k = rng.normal(size=(80, d)) * 1.5 if general else GetKs(cfg, getK(chi, 1.0, d))
    v = 0.5 + rng.random(len(k))
    w = 0.2 + 1.6 * rng.random(N)
    base = ShiftedCCPotential(d)
    for kk, vv in zip(k, v):
        base.AddConstraint(kk, [vv])
    base.Shift = 0.125
    base.SetRhoBackend("nufft")  # a weighted potential keeps the direct sums
    pot = ShiftedCCPotential_withWeight(base, w)
    pot.SetConfiguration(cfg)
    assert pot.GetLastRhoBackendUsed() == "direct"
    assert pot.UsesLatticePath() == (not general)
    pot.GetRho()
    E, F = pot.EnergyAndAllForce()
    re, im = po.rho_weighted(cfg.relative, cfg.basis, k, w)
    vol = po.volume(cfg.basis)
    Eo = po.energy(v, re, im, vol, 0.125)
    Fo = po.weighted_forces(cfg.relative, cfg.basis, k, v, w, re, im)
    W = np.sum(np.abs(w))
    kmax = np.max(np.linalg.norm(k, axis=1))
    assert np.max(np.abs(pot.RhoReal - re)) <= RTOL * W and np.max(np.abs(pot.RhoImag - im)) <= RTOL * W
    assert abs(E - Eo) <= RTOL * np.sum(v) * W * W / vol
    assert np.max(np.abs(F - Fo)) <= RTOL * 4 * np.sum(v) * W * kmax * np.max(w) / vol
    with pytest.raises(CCGPUError):  # one weight per particle
        pot.SetConfiguration(RandomConfiguration(seed, N + 1, d))
        pot.Energy()


def test_log_form_matches_oracle_and_relaxes():
    from modules_amd import LogCCPotential, ShiftedCCPotential
    cfg = RandomConfiguration(12, 300, 2)
    k = GetKs(cfg, getK(0.3, 1.0, 2))
    v = np.ones(len(k))
    base = ShiftedCCPotential(2)
    for kk in k:
        base.AddConstraint(kk, [1.0])
    base.Shift = 0.5
    pot = LogCCPotential(base)
    pot.SetConfiguration(cfg)
    E, F = pot.EnergyAndAllForce()
    re, im = po.rho(cfg.relative, cfg.basis, k)
    Ep = po.energy(v, re, im, po.volume(cfg.basis), 0.5)
    Fp = po.forces(cfg.relative, cfg.basis, k, v, re, im)
    assert E == pytest.approx(np.log(Ep), rel=1e-13)
    assert np.allclose(F, Fp / Ep, rtol=1e-11, atol=1e-13 * np.max(np.abs(Fp / Ep)))
    c1 = cfg.copy()
    H, ev, _ = RelaxStructure_NLOPT(c1, pot, 0.0, 0, 0.0, 1500)
    re, im = po.rho(c1.relative, c1.basis, k)
    assert H == pytest.approx(np.log(po.energy(v, re, im, po.volume(c1.basis), 0.5)), rel=1e-10)
    assert H < E and ev > 0


# ---------------------------------------------------------------------------------------------------------------
# sub-configurations (ShiftedCCPotential_S2_MultiConfig)
# ---------------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("nsub,Np,d,chi,seed,general", [(4, 300, 2, 0.3, 1, False), (3, 500, 3, 0.35, 2, False),
                                                        (5, 120, 2, 0.3, 3, True), (1, 400, 2, 0.3, 4, False)])
def test_s2_multiconfig_matches_oracle(nsub, Np, d, chi, seed, general):
    from modules_amd import ShiftedCCPotential_S2_MultiConfig
    N = nsub * Np
    cfg = RandomConfiguration(seed, N, d)
    rng = np.random.default_rng(seed)
    k = rng.normal(size=(70, d)) * 1.5 if general else GetKs(cfg, getK(chi, 1.0, d))
    v = 0.5 + rng.random(len(k))
    s0 = rng.random(len(k)) * 2
    pot = ShiftedCCPotential_S2_MultiConfig(d, nsub)
    for kk, vv, ss in zip(k, v, s0):
        pot.AddConstraint(kk, [vv, ss])
    pot.Shift = 0.3
    pot.SetConfiguration(cfg)
    assert pot.UsesLatticePath() == (not general)
    pot.GetRho()
    E, F = pot.EnergyAndAllForce()
    Eo, Fo, re, im = po.s2_multiconfig(cfg.relative, cfg.basis, k, v, s0, nsub, 0.3)
    assert np.max(np.abs(pot.RhoReal.reshape(nsub, -1) - re)) <= RTOL * Np
    assert np.max(np.abs(pot.RhoImag.reshape(nsub, -1) - im)) <= RTOL * Np
    snow = (re ** 2 + im ** 2).sum(axis=0) / N
    assert np.allclose(pot.Snow, snow, rtol=1e-12, atol=1e-12)
    _, se, sf = s2_scales(k, v, s0, Np)
    assert abs(E - Eo) <= RTOL * se
    assert np.max(np.abs(F - Fo)) <= RTOL * sf
    for i in (0, N // 2, N - 1):
        assert np.allclose(pot.Force(i), Fo[i], rtol=0, atol=RTOL * sf)


def test_s2_multiconfig_requires_a_multiple_and_relaxes():
    from modules_amd import ShiftedCCPotential_S2_MultiConfig
    pot = ShiftedCCPotential_S2_MultiConfig(2, 3)
    cfg = RandomConfiguration(6, 150, 2)
    k = GetKs(cfg, getK(0.25, 1.0, 2))
    for kk in k:
        pot.AddConstraint(kk, [1.0, 0.5])
    with pytest.raises(CCGPUError):
        pot.SetConfiguration(RandomConfiguration(6, 151, 2))
    pot.SetConfiguration(cfg)
    E0 = pot.Energy()
    c1 = cfg.copy()
    E1, ev, _ = RelaxStructure_NLOPT(c1, pot, 0.0, 0, 0.0, 1500)
    Eo = po.s2_multiconfig(c1.relative, c1.basis, k, np.ones(len(k)), np.full(len(k), 0.5), 3)[0]
    assert E1 == pytest.approx(Eo, rel=1e-10, abs=1e-14)
    assert E1 < 0.01 * E0 and ev > 0


# ---------------------------------------------------------------------------------------------------------------
# Monte Carlo trial moves (ShiftedCCPotential_ForMCMixin)
# ---------------------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("d,N,chi,s2,general", [(2, 300, 0.3, False, False), (3, 400, 0.35, False, False),
                                                (2, 200, 0.3, True, False), (2, 150, 0.3, False, True)])
def test_mc_moves_match_recomputed_energies(d, N, chi, s2, general):
    from modules_amd import ShiftedCCPotential, ShiftedCCPotential_ForMCMixin
    cfg = RandomConfiguration(40 + d, N, d)
    rng = np.random.default_rng(d + N)
    k = rng.normal(size=(60, d)) * 1.4 if general else GetKs(cfg, getK(chi, 1.0, d))
    v = 0.5 + rng.random(len(k))
    t = rng.random(len(k))
    base = ShiftedCCPotential_S2(d) if s2 else ShiftedCCPotential(d)
    for i in range(len(k)):
        base.AddConstraint(k[i], [v[i], t[i]] if s2 else [v[i]])
    base.Shift = 0.25
    pot = ShiftedCCPotential_ForMCMixin(base)
    pot.SetConfiguration(cfg)

    def oracle_E(c):
        re, im = po.rho(c.relative, c.basis, k)
        if s2:
            return po.s2_energy(v, t, re, im, N, 0.25), re, im
        return po.energy(v, re, im, po.volume(c.basis), 0.25), re, im

    E = pot.Energy()
    accepted = 0
    for step in range(40):
        i = int(rng.integers(N))
        prev = cfg.GetCartesianCoordinates(i).copy()
        new_rel = cfg.relative[i] + 0.05 * rng.normal(size=d)
        trial = cfg.copy()
        trial.MoveParticle(i, new_rel)
        after = trial.GetCartesianCoordinates(i)
        dE = pot.TryMove(prev, after, i if step % 2 == 0 else -1)
        Et = oracle_E(trial)[0]
        assert dE == pytest.approx(Et - E, rel=1e-9, abs=1e-10 * max(1.0, abs(E)))
        if dE < 0 or rng.random() < 0.3:
            pot.AcceptMove()
            cfg.MoveParticle(i, new_rel)
            E = E + dE
            accepted += 1
            assert pot.Energy() == pytest.approx(E, rel=1e-12, abs=1e-14)
            if step % 2 == 1:  # accepted without its index: bind again before forces
                with pytest.raises(CCGPUError):
                    pot.AllForce()
                pot.SetConfiguration(cfg)
                E = pot.Energy()
    assert accepted > 5
    Eo, re, im = oracle_E(cfg)
    assert pot.Energy() == pytest.approx(Eo, rel=1e-10, abs=1e-12)
    pot.GetRho()
    assert np.max(np.abs(pot.RhoReal - re)) <= 1e-10 * N and np.max(np.abs(pot.RhoImag - im)) <= 1e-10 * N
    # forces after the last accepted move (index given or configuration rebound) follow the moved particles
    F = pot.AllForce()
    Fo = po.s2_forces(cfg.relative, cfg.basis, k, v, t, re, im) if s2 else \
        po.forces(cfg.relative, cfg.basis, k, v, re, im)
    assert np.max(np.abs(F - Fo)) <= 1e-9 * max(1.0, np.max(np.abs(Fo)))


def test_indexed_view_equals_the_sub_configuration():
    """SetIndexedConfiguration(full, indices) gives the bits of SetConfiguration on the sub-configuration
    (CollectiveCoordinatePotential.h:1477-1496), and refuses what the reference refuses."""
    from modules_amd import ShiftedCCPotential
    full = RandomConfiguration(31, 900, 3)
    k = GetKs(full, getK(0.3, 1.0, 3))
    idx = np.random.default_rng(2).permutation(900)[:333]
    a, b = ShiftedCCPotential(3), ShiftedCCPotential(3)
    for kk in k:
        a.AddConstraint(kk, [1.0])
        b.AddConstraint(kk, [1.0])
    a.SetIndexedConfiguration(full, idx)
    b.SetConfiguration(Configuration(full.basis, full.relative[idx]))
    Ea, Fa = a.EnergyAndAllForce()
    Eb, Fb = b.EnergyAndAllForce()
    assert Ea == Eb and np.array_equal(Fa, Fb)
    with pytest.raises(IndexError):
        a.SetIndexedConfiguration(full, [0, 900])
    s2 = ShiftedCCPotential_S2(3)
    with pytest.raises(RuntimeError):
        s2.SetIndexedConfiguration(full, idx)
