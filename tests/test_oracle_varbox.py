"""Pins of the oracle's target-S(k) and variable-cell restatements (CPU only).

  * S2 energy / forces (CollectiveCoordinatePotential.h:1686-1741): forces against central differences of the
    energy; the energy against a closed form on a perfect sublattice (rho = N on its reciprocal lattice, 0 off it)
  * the weight / target catalog (CCPotential_Kernels.h:18-340): values at chosen k, gradients against central
    differences (this pins the lemniscate's hand-derived gradient too)
  * the k-point list (CCPotential_varBox.h:129-208): the count for a cubic cell against the sum-of-three-squares
    counts r_3(n), and the set against the oracle's GetKs (a different reference routine) for the same radius
  * the Shift and EnergyDerivativeToBasisVectors of the three forms: against central differences of
    E + P V over every basis component at fixed relative coordinates (the reference's own check,
    potential/Potential.cpp:21-40), on skewed cells, so a dropped -Phi/V term, a sign or a transposed
    inverse fails
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from oracle import varbox as vb


def test_s2_forces_are_minus_the_energy_gradient():
    rng = np.random.default_rng(3)
    a = np.array([[4.0, 0.0], [0.7, 3.5]])
    s = rng.random((7, 2))
    k, _ = po.get_ks(a, 4.0)
    v = 0.5 + rng.random(len(k))
    s0 = rng.random(len(k))
    N = len(s)

    def E(ss):
        re, im = po.rho(ss, a, k)
        return po.s2_energy(v, s0, re, im, N, 0.3)

    re, im = po.rho(s, a, k)
    F = po.s2_forces(s, a, k, v, s0, re, im)
    h = 1e-6
    inv = np.linalg.inv(a)  # r = s a  ->  ds = dr a^{-1}
    for j in range(N):
        for c in range(2):
            dr = np.zeros(2)
            dr[c] = h
            sp, sm = s.copy(), s.copy()
            sp[j] += dr @ inv
            sm[j] -= dr @ inv
            fd = -(E(sp) - E(sm)) / (2 * h)
            assert abs(fd - F[j, c]) < 1e-6 * max(1.0, abs(F[j, c]))


def test_s2_energy_closed_form_on_a_sublattice():
    # 3 x 3 particles on a square sublattice of an L = 6 cell: rho(k) = 9 where k is on the sublattice's reciprocal
    # lattice (n multiple of 3), 0 elsewhere; S = |rho|^2 / N = 9 or 0
    L = 6.0
    a = np.diag([L, L])
    s = np.array([[i / 3.0, j / 3.0] for i in range(3) for j in range(3)])
    k = []
    n = []
    for x in range(-4, 5):
        for y in range(-4, 5):
            if (x, y) > (0, 0) and x * x + y * y <= 20:
                k.append([2 * np.pi * x / L, 2 * np.pi * y / L])
                n.append((x, y))
    k = np.array(k)
    v = np.linspace(0.5, 1.5, len(k))
    s0 = np.full(len(k), 0.25)
    re, im = po.rho(s, a, k)
    S = np.array([9.0 if (x % 3 == 0 and y % 3 == 0) else 0.0 for x, y in n])
    assert np.allclose((re ** 2 + im ** 2) / 9, S, atol=1e-12)
    want = float(np.sum(v * (S - 0.25) ** 2)) + 1.5
    assert po.s2_energy(v, s0, re, im, 9, 1.5) == pytest.approx(want, rel=1e-13)
    # on the plain form the same rho gives (1/V) sum v |rho|^2
    assert po.energy(v, re, im, L * L) == pytest.approx(float(np.sum(v * 9 * S)) / (L * L), rel=1e-13)


@pytest.mark.parametrize("kind,k,want", [
    ("K2", [0.3, 0.4], 0.25), ("KM2", [0.3, 0.4], 1.0), ("KM3", [0.3, 0.4], 1.0), ("KM4", [0.3, 0.4], 1.0),
    ("KM2", [0.0, 0.25], 9.0), ("KM3", [0.0, 0.25], 27.0), ("K2", [0.6, 0.8], 0.0)])
def test_weights_at_chosen_points(kind, k, want):
    assert vb.vk(kind, k) == pytest.approx(want, rel=1e-14, abs=1e-15)


@pytest.mark.parametrize("kind,k,n,want", [
    ("K2", [0.3, 0.4], 1, 0.25), ("Kn", [0.3, 0.4], 3, 0.125), ("Klc", [0.3, 0.4], 2, 0.625),
    ("KM2", [0.3, 0.4], 1, 1.0), ("KMLog", [0.3, 0.4], 1, np.log(2.0))])
def test_targets_at_chosen_points(kind, k, n, want):
    assert vb.s0k(kind, k, n) == pytest.approx(want, rel=1e-14)


def _fd_grad(f, k, h=1e-6):
    k = np.asarray(k, dtype=np.float64)
    g = np.zeros_like(k)
    for c in range(len(k)):
        e = np.zeros_like(k)
        e[c] = h
        g[c] = (f(k + e) - f(k - e)) / (2 * h)
    return g


@pytest.mark.parametrize("kind", ["K2", "KM2", "KM3", "KM4"])
@pytest.mark.parametrize("k", [[0.3, 0.5], [0.2, -0.1, 0.6]])
def test_weight_gradients(kind, k):
    assert np.allclose(vb.grad_vk(kind, k), _fd_grad(lambda q: vb.vk(kind, q), k), rtol=1e-7, atol=1e-7)


@pytest.mark.parametrize("k", [[0.8, 0.2], [0.5, -0.3], [0.9, 0.05]])
def test_lemniscate_gradient(k):
    Rc = 1.3
    assert vb.vk("Lemniscate", k, Rc) > 0
    assert np.allclose(vb.grad_vk("Lemniscate", k, Rc), _fd_grad(lambda q: vb.vk("Lemniscate", q, Rc), k),
                       rtol=1e-6, atol=1e-9)
    # outside the lemniscate (|kx| < |ky|) the weight and its gradient vanish
    assert vb.vk("Lemniscate", [0.1, 0.5], Rc) == 0.0 and not np.any(vb.grad_vk("Lemniscate", [0.1, 0.5], Rc))


@pytest.mark.parametrize("kind,n", [("K2", 1), ("Kn", 2.5), ("Klc", 1.5), ("KM2", 1), ("KMLog", 1)])
@pytest.mark.parametrize("k", [[0.3, 0.5], [0.2, -0.1, 0.6]])
def test_target_gradients(kind, n, k):
    assert np.allclose(vb.grad_s0k(kind, k, n), _fd_grad(lambda q: vb.s0k(kind, q, n), k), rtol=1e-7, atol=1e-7)


def test_kpoint_count_against_sums_of_three_squares():
    # cubic cell L = 2 pi: k = n, 0 < |n|^2 < 6.25 -> r_3(1..6) = 6, 12, 8, 6, 24, 24 -> 80 points, 40 pairs
    L = 2 * np.pi
    k, n = vb.kpoints(np.diag([L, L, L]), 2.5)
    assert len(k) == 40
    assert np.allclose(k, n)
    for row in n:
        nz = [x for x in row if x != 0]
        assert nz[0] > 0
    assert len({tuple(r) for r in n} | {tuple(-r) for r in n}) == 80
    # 2D: r_2(1..4) = 4, 4, 0, 4 below |n|^2 < 4.5 -> 12 points, 6 pairs
    assert len(vb.kpoints(np.diag([L, L]), 2.12)[0]) == 6


@pytest.mark.parametrize("a,rK", [(np.array([[5.0, 0.0], [1.2, 4.0]]), 4.1),
                                  (np.array([[4.0, 0, 0], [0.5, 4.5, 0], [0.3, -0.6, 5.0]]), 3.3)])
def test_kpoints_match_getks_for_the_same_radius(a, rK):
    k, n = vb.kpoints(a, rK)
    kg, _ = po.get_ks(a, rK)
    key = lambda arr: sorted(tuple(np.round(r, 9)) for r in arr)  # noqa: E731
    # GetKs keeps one of each pair too; compare as sets of +-pairs (the half-space conventions may differ)
    sym = lambda arr: sorted(set(key(arr)) | set(key(-np.asarray(arr))))  # noqa: E731
    assert len(k) == len(kg)
    assert sym(k) == sym(kg)
    assert np.all(np.linalg.norm(k, axis=1) < rK)


CASES = [
    vb.Spec("stealthy", "K2", None, 1.0),
    vb.Spec("stealthy", "KM3", None, 1.0),
    vb.Spec("S2", "K2", "K2", 1.0),
    vb.Spec("S2", "KM3", "Kn", 1.0, 1.5),
    vb.Spec("S2", "KM2", "KMLog", 1.0),
    vb.Spec("S2_noV0", "K2", "K2", 1.0),
    vb.Spec("S2_noV0", "KM3", "Klc", 1.0, 2.0),
]


@pytest.mark.parametrize("spec", CASES, ids=lambda s: f"{s.form}-{s.vk}-{s.s0}")
@pytest.mark.parametrize("d", [2, 3])
def test_cell_derivative_against_finite_differences(spec, d):
    rng = np.random.default_rng(11 + d)
    N = 6 if d == 2 else 5
    s = rng.random((N, d))
    if d == 2:
        a = np.array([[9.0, 0.0], [1.3, 8.2]])
    else:
        a = np.array([[6.5, 0.0, 0.0], [0.7, 6.1, 0.0], [-0.4, 0.9, 6.8]])
    P = 0.37
    g = vb.derivative(spec, a, s, P)
    h = 1e-6
    for idx in range(d * d):
        ap, am = a.copy(), a.copy()
        ap[idx // d, idx % d] += h
        am[idx // d, idx % d] -= h
        Hp = vb.energy(spec, ap, s)[0] + P * po.volume(ap)
        Hm = vb.energy(spec, am, s)[0] + P * po.volume(am)
        fd = (Hp - Hm) / (2 * h)
        assert abs(fd - g[idx]) <= 2e-6 * max(1.0, abs(g[idx])), (idx, fd, g[idx])


def test_shift_closed_forms():
    v = np.array([0.5, 0.25, 0.125])
    assert vb.shift("stealthy", 10, 20.0, v) == pytest.approx(10 * 9 / 20 / 2 - 10 / 20 * 0.875)
    assert vb.shift("S2", 10, 20.0, v, np.ones(3)) == 0.0
    assert vb.shift("S2_noV0", 10, 20.0, v, np.array([0.0, 2.0, 3.0])) == pytest.approx(-(0.5 + 0.25 + 0.5))


def test_weighted_rho_pins():
    """ShiftedCCPotential_withWeight's rho: unit weights give the plain rho; integer weights equal duplicating the
    particles; forces are minus the gradient of the energy over the weighted rho."""
    rng = np.random.default_rng(5)
    a = np.array([[5.0, 0.0], [0.8, 4.5]])
    s = rng.random((9, 2))
    k, _ = po.get_ks(a, 3.5)
    re1, im1 = po.rho_weighted(s, a, k, np.ones(9))
    re0, im0 = po.rho(s, a, k)
    assert np.array_equal(re1, re0) and np.array_equal(im1, im0)
    w = np.array([1, 2, 3, 1, 1, 2, 1, 4, 1], dtype=float)
    rew, imw = po.rho_weighted(s, a, k, w)
    dup = np.repeat(s, w.astype(int), axis=0)
    red, imd = po.rho(dup, a, k)
    assert np.allclose(rew, red, atol=1e-12) and np.allclose(imw, imd, atol=1e-12)
    v = 0.5 + rng.random(len(k))
    vol = po.volume(a)
    wr = 0.3 + rng.random(9)
    re, im = po.rho_weighted(s, a, k, wr)
    F = po.weighted_forces(s, a, k, v, wr, re, im)
    inv = np.linalg.inv(a)
    h = 1e-6
    for j in range(9):
        for c in range(2):
            dr = np.zeros(2)
            dr[c] = h
            sp, sm = s.copy(), s.copy()
            sp[j] += dr @ inv
            sm[j] -= dr @ inv
            Ep = po.energy(v, *po.rho_weighted(sp, a, k, wr), vol)
            Em = po.energy(v, *po.rho_weighted(sm, a, k, wr), vol)
            assert abs(-(Ep - Em) / (2 * h) - F[j, c]) < 1e-6 * max(1.0, abs(F[j, c]))


def test_s2_multiconfig_pins():
    """ShiftedCCPotential_S2_MultiConfig: one sub-configuration is the plain S2 form; C identical sub-configurations
    give the S2 energy of one of them with forces divided by C (the mean S equals each S); forces are minus the
    energy gradient."""
    rng = np.random.default_rng(9)
    a = np.array([[6.0, 0.0], [0.9, 5.5]])
    s = rng.random((8, 2))
    k, _ = po.get_ks(a, 3.0)
    v = 0.5 + rng.random(len(k))
    s0 = rng.random(len(k))
    E1, F1, re, im = po.s2_multiconfig(s, a, k, v, s0, 1, 0.2)
    r0, i0 = po.rho(s, a, k)
    assert E1 == pytest.approx(po.s2_energy(v, s0, r0, i0, 8, 0.2), rel=1e-14)
    assert np.allclose(F1, po.s2_forces(s, a, k, v, s0, r0, i0), rtol=1e-13, atol=1e-15)
    E3, F3, re3, im3 = po.s2_multiconfig(np.vstack([s, s, s]), a, k, v, s0, 3, 0.2)
    assert E3 == pytest.approx(E1, rel=1e-12)
    assert np.allclose(F3, np.vstack([F1, F1, F1]) / 3, rtol=1e-12, atol=1e-15)
    s6 = rng.random((12, 2))
    E, F, _, _ = po.s2_multiconfig(s6, a, k, v, s0, 2)
    inv = np.linalg.inv(a)
    h = 1e-6
    for j in range(12):
        for c in range(2):
            dr = np.zeros(2)
            dr[c] = h
            sp, sm = s6.copy(), s6.copy()
            sp[j] += dr @ inv
            sm[j] -= dr @ inv
            fd = -(po.s2_multiconfig(sp, a, k, v, s0, 2)[0] - po.s2_multiconfig(sm, a, k, v, s0, 2)[0]) / (2 * h)
            assert abs(fd - F[j, c]) < 1e-6 * max(1.0, abs(F[j, c]))
