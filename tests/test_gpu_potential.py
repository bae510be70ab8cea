"""GPU parity of rho / energy / forces against the oracle (the reference direct path restated in C).

Tolerances follow the reference test's scale-aware closeness (potential/tests/finufft_type1/
test_finufft_type1.cpp:197-212): rho ~ N, energy ~ sum|V| N^2 / vol, force ~ 4 sum|V| N kmax / vol,
with relative tolerance 1e-12 on those scales (the parity bar: FP64 within 1e-12 relative).
"""
import numpy as np
import pytest

from modules_amd import Configuration, GetKs, RandomConfiguration, ShiftedCCPotential, getK
from modules_amd._lib import CCGPUError
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def scales(k, v, N, vol):
    sumV = np.sum(np.abs(v))
    kmax = np.max(np.linalg.norm(k, axis=1)) if len(k) else 0.0
    return max(N, 1.0), sumV * N * N / vol, 4.0 * sumV * N * kmax / vol


def make_pot(d, k, v, shift=0.0):
    pot = ShiftedCCPotential(d)
    for kk, vv in zip(k, v):
        pot.AddConstraint(kk, [vv])
    pot.Shift = shift
    return pot


def compare(cfg, k, v, shift=0.0, tiles=None, expect_lattice=None):
    d = cfg.GetDimension()
    N = cfg.NumParticle()
    pot = make_pot(d, k, v, shift)
    if tiles:
        pot.SetConfiguration(cfg)
        pot.SetTiles(*tiles)
    pot.SetConfiguration(cfg)
    pot.GetRho()
    E, F = pot.EnergyAndAllForce()
    re, im = po.rho(cfg.relative, cfg.basis, k)
    vol = po.volume(cfg.basis)
    Eo = po.energy(v, re, im, vol, shift)
    Fo = po.forces(cfg.relative, cfg.basis, k, v, re, im)
    sr, se, sf = scales(k, v, N, vol)
    if len(k):
        assert np.max(np.abs(pot.RhoReal - re)) <= RTOL * sr
        assert np.max(np.abs(pot.RhoImag - im)) <= RTOL * sr
    assert abs(E - Eo) <= RTOL * max(se, abs(shift), 1e-300) + 1e-300
    if N:
        assert np.max(np.abs(F - Fo)) <= RTOL * max(sf, 1e-300)
    if expect_lattice is not None:
        assert pot.UsesLatticePath() == expect_lattice
    return pot, E, F


@pytest.mark.parametrize("N,d,chi,seed", [(200, 2, 0.3, 0), (500, 2, 0.3, 7), (2000, 3, 0.45, 1), (300, 3, 0.2, 3)])
def test_stealthy_configs_match_oracle(N, d, chi, seed):
    cfg = RandomConfiguration(seed, N, d)
    k = GetKs(cfg, getK(chi, 1.0, d))
    compare(cfg, k, np.ones(len(k)), expect_lattice=True)


@pytest.mark.parametrize("seg", [4, 8, 16, 24, 32])
@pytest.mark.parametrize("rho_kernel", [0, 1, 2])
def test_segment_lengths_agree(seg, rho_kernel):
    cfg = RandomConfiguration(11, 700, 2)
    k = GetKs(cfg, getK(0.4, 1.0, 2))
    pot = make_pot(2, k, np.ones(len(k)))
    pot.SetRhoKernel(rho_kernel)
    pot.SetConfiguration(cfg)
    pot.SetTiles(seg, 96, 7)
    pot.GetRho()
    E, F = pot.EnergyAndAllForce()
    re, im = po.rho(cfg.relative, cfg.basis, k)
    vol = po.volume(cfg.basis)
    sr, se, sf = scales(k, np.ones(len(k)), 700, vol)
    assert np.max(np.abs(pot.RhoReal - re)) <= RTOL * sr and np.max(np.abs(pot.RhoImag - im)) <= RTOL * sr
    assert abs(E - po.energy(np.ones(len(k)), re, im, vol)) <= RTOL * se
    assert np.max(np.abs(F - po.forces(cfg.relative, cfg.basis, k, np.ones(len(k)), re, im))) <= RTOL * sf


@pytest.mark.parametrize("seg", [16, 24, 32])
def test_phase_recurrences_at_degenerate_step_angles(seg):
    # The lattice kernels walk each segment by the three-term recurrence e_{t+1} = 2 cos(phi) e_t - e_{t-1}
    # (phi = 2 pi s_last).  Its error amplification is sin(n phi) / sin(phi) <= n, largest where sin(phi) -> 0:
    # particles at s_last = 0, 1e-12, 1/2 -+ 1e-12, 1 - 1e-12 and 1/4, with segments as long as the tile allows
    # (N = 64 in an 8 x 8 cell, |k| < 17: lattice lines of about 44 points), must still meet the oracle.
    rng = np.random.default_rng(5)
    s = rng.random((64, 2))
    s[:36, 1] = np.repeat([0.0, 1e-12, 0.5 - 1e-12, 0.5, 1.0 - 1e-12, 0.25], 6)
    cfg = Configuration(np.eye(2) * 8.0, s)
    k = GetKs(cfg, 17.0)
    assert len(k) > 700
    pot, _, _ = compare(cfg, k, np.ones(len(k)), tiles=(seg, 64, 5), expect_lattice=True)
    assert pot.UsesLatticePath()


def test_general_path_random_wavevectors():
    rng = np.random.default_rng(5)
    cfg = Configuration(np.array([[3.0, 0.0, 0.0], [0.5, 2.5, 0.0], [0.2, 0.3, 4.0]]), rng.random((150, 3)))
    k = rng.normal(size=(90, 3)) * 2.0
    v = rng.uniform(-0.5, 1.5, size=90)
    compare(cfg, k, v, shift=0.3, expect_lattice=False)


def test_skew_cell_lattice_modes_with_duplicates_and_zero():
    rng = np.random.default_rng(606)
    basis = np.array([[1.3, 0.0], [0.45, 0.9]]) * 4
    cfg = Configuration(basis, rng.random((30, 2)))
    b = po.reciprocal_basis(basis)
    modes = [(m0, m1) for m1 in range(-4, 5) for m0 in range(-4, 5)]
    modes += modes[:5]  # duplicates
    rng.shuffle(modes)
    k = np.array([m0 * b[0] + m1 * b[1] for m0, m1 in modes])
    v = rng.uniform(0.2, 1.2, size=len(k))
    compare(cfg, k, v, expect_lattice=True)


def test_sign_analytic_phases():
    # TestSign (test_finufft_type1.cpp:387-439): 1D s=0.25 gives rho(+1)=+i, rho(-1)=-i, rho(0)=N
    cfg = Configuration(np.array([[1.0]]), np.array([[0.25]]))
    b = po.reciprocal_basis(cfg.basis)
    pot = make_pot(1, np.array([[b[0, 0]], [-b[0, 0]], [0.0]]), [1.0, 1.0, 1.0])
    pot.SetConfiguration(cfg)
    pot.GetRho()
    assert abs(pot.RhoReal[0]) <= 1e-10 and abs(pot.RhoImag[0] - 1.0) <= 1e-10
    assert abs(pot.RhoImag[1] + 1.0) <= 1e-10
    assert abs(pot.RhoReal[2] - 1.0) <= 1e-10
    # 2D skew cell: cos/sin of 2 pi (m . s)
    basis = np.array([[1.0, 0.0], [0.4, 0.9]])
    cfg2 = Configuration(basis, np.array([[0.25, 0.125]]))
    b2 = po.reciprocal_basis(basis)
    k = np.array([1 * b2[0] + 1 * b2[1], -2 * b2[0] + 3 * b2[1]])
    pot2 = make_pot(2, k, [1.0, 1.0])
    pot2.SetConfiguration(cfg2)
    pot2.GetRho()
    p0 = 2 * np.pi * (0.25 + 0.125)
    p1 = 2 * np.pi * (-2 * 0.25 + 3 * 0.125)
    assert abs(pot2.RhoReal[0] - np.cos(p0)) <= 1e-10 and abs(pot2.RhoImag[0] - np.sin(p0)) <= 1e-10
    assert abs(pot2.RhoReal[1] - np.cos(p1)) <= 1e-10 and abs(pot2.RhoImag[1] - np.sin(p1)) <= 1e-10


def test_empty_constraints_and_particles():
    # TestEmpty (test_finufft_type1.cpp:442-486)
    rng = np.random.default_rng(202)
    cfg = Configuration(np.array([[1.0, 0.0], [0.3, 1.1]]), rng.random((12, 2)))
    pot = ShiftedCCPotential(2)
    pot.Shift = 0.7
    pot.SetConfiguration(cfg)
    pot.GetRho()
    assert pot.RhoReal.size == 0
    assert pot.Energy() == 0.7
    F = pot.AllForce()
    assert F.shape == (12, 2) and np.all(F == 0.0)
    cfg3 = Configuration(np.eye(3), np.zeros((0, 3)))
    b = po.reciprocal_basis(np.eye(3))
    modes = [(i, j, l) for l in (-1, 0, 1) for j in (-1, 0, 1) for i in (-1, 0, 1)]
    k = np.array([i * b[0] + j * b[1] + l * b[2] for i, j, l in modes])
    v = rng.uniform(0.5, 1.5, size=27)
    pot3 = make_pot(3, k, v)
    pot3.SetConfiguration(cfg3)
    pot3.GetRho()
    assert np.all(pot3.RhoReal == 0.0) and np.all(pot3.RhoImag == 0.0)
    assert pot3.Energy() == 0.0
    assert pot3.AllForce().shape == (0, 3)


def test_finite_difference_forces():
    # TestFiniteDifference (test_finufft_type1.cpp:660-702): central differences, h = 5e-5
    rng = np.random.default_rng(606)
    basis = np.array([[1.3, 0.0], [0.45, 0.9]])
    cfg = Configuration(basis, rng.random((30, 2)))
    b = po.reciprocal_basis(basis)
    modes = [(m0, m1) for m1 in range(-4, 5) for m0 in range(-4, 5)]
    k = np.array([m0 * b[0] + m1 * b[1] for m0, m1 in modes])
    v = rng.uniform(0.2, 1.2, size=len(k))
    pot = make_pot(2, k, v)
    pot.SetConfiguration(cfg)
    F = pot.AllForce()
    _, _, sf = scales(k, v, 30, po.volume(basis))
    h = 5e-5
    inv = np.linalg.inv(basis)
    for i in (0, 17):
        for dd in range(2):
            E = []
            for sgn in (+1, -1):
                c = cfg.copy()
                r = c.GetCartesianCoordinates(i)
                r[dd] += sgn * h
                c.MoveParticle(i, r @ inv)
                pot.SetConfiguration(c)
                E.append(pot.Energy())
            fd = -(E[0] - E[1]) / (2 * h)
            assert abs(F[i, dd] - fd) <= 1e-5 * sf + 1e-8


def test_momentum_sum_vanishes():
    cfg = RandomConfiguration(3, 400, 3)
    k = GetKs(cfg, getK(0.3, 1.0, 3))
    _, _, F = compare(cfg, k, np.ones(len(k)))
    assert np.max(np.abs(F.sum(axis=0))) <= 1e-10 * np.max(np.abs(F)) * 400


def test_repeatable_bits():
    cfg = RandomConfiguration(9, 1000, 2)
    k = GetKs(cfg, getK(0.35, 1.0, 2))
    p1, E1, F1 = compare(cfg, k, np.ones(len(k)))
    p1.SetConfiguration(cfg)
    E2, F2 = p1.EnergyAndAllForce()
    assert E1 == E2 and np.array_equal(F1, F2)


def test_errors_are_raised():
    pot = ShiftedCCPotential(2)
    with pytest.raises(RuntimeError):
        pot.Energy()
    with pytest.raises(ValueError):
        pot.AddConstraint([1.0, 0.0], [])
    cfg = RandomConfiguration(1, 10, 2)
    pot.SetConfiguration(cfg)
    with pytest.raises(CCGPUError):
        pot.Force(10)
    bad = cfg.copy()
    bad.relative[0, 0] = np.nan
    with pytest.raises(CCGPUError):
        pot.SetConfiguration(bad)


def test_autotuned_tiles_are_used_and_keep_parity():
    from modules_amd import autotune
    autotune.clear()
    best = autotune.autotune(600, 3, chi=0.3, batch=4, repeats=2)
    assert best["segment_len"] in (8, 16, 24, 32) and best["ms_per_batched_ef"] > 0
    cfg = RandomConfiguration(17, 600, 3)
    k = GetKs(cfg, getK(0.3, 1.0, 3))
    compare(cfg, k, np.ones(len(k)), expect_lattice=True)
    autotune.clear()


@pytest.mark.parametrize("N,d,chi,seed", [(2000, 2, 0.4, 4), (2000, 3, 0.45, 1), (700, 2, 0.3, 11)])
def test_mixed_precision_error_against_oracle(N, d, chi, seed):
    # mixed mode: rho (and so the energy and the force coefficients) in FP64; the force pass reduces each phase in
    # FP64, then runs the sincos, the phase chains and the force terms in FP32 and sums segments in FP64.  FP32 has
    # a unit roundoff of 6e-8; per-term errors of a few 1e-6 (the chains' t^2 u growth) summed over the constraints
    # bound the force error at 1e-5 of the scale (F ~ 4 sum|V| N kmax / vol); rho and E meet the FP64 bar
    cfg = RandomConfiguration(seed, N, d)
    k = GetKs(cfg, getK(chi, 1.0, d))
    v = np.ones(len(k))
    pot = make_pot(d, k, v)
    pot.SetPrecision("mixed")
    pot.SetConfiguration(cfg)
    pot.GetRho()
    E, F = pot.EnergyAndAllForce()
    re, im = po.rho(cfg.relative, cfg.basis, k)
    vol = po.volume(cfg.basis)
    Eo = po.energy(v, re, im, vol)
    Fo = po.forces(cfg.relative, cfg.basis, k, v, re, im)
    sr, se, sf = scales(k, v, N, vol)
    drho = max(np.max(np.abs(pot.RhoReal - re)), np.max(np.abs(pot.RhoImag - im))) / sr
    dE = abs(E - Eo) / se
    dF = np.max(np.abs(F - Fo)) / sf
    print(f"mixed N={N} d={d}: rho {drho:.2e}, E {dE:.2e}, F {dF:.2e} (relative to the scales)")
    assert drho <= RTOL and dE <= RTOL and dF <= 1e-5
    # and the force pass is not the FP64 one: the FP32 rounding shows
    assert dF > 1e-13
    pot.SetPrecision("fp64")
    pot.SetConfiguration(cfg)
    E64, F64 = pot.EnergyAndAllForce()
    assert abs(E64 - Eo) <= RTOL * se
