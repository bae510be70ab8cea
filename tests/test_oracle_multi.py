"""Pins of the oracle's multi-species potential (MultiStealthyPotential, apps/CCOptimization/
MultiStealthyPotential.cpp:682-833) against what the mathematics fixes:
  * one species holding every particle, equal radii r: the single-species oracle with sigma = 2r (itself pinned in
    test_oracle.py), energy and forces;
  * F = -grad E by central differences, several species, unequal radii, a skewed cell;
  * momentum conservation (sum of forces zero) and invariance under relabelling the particles;
  * two particles in a large box: the closed form e0 (1 - r/(r1+r2))^2 and its force;
  * contacts through several periodic images (a tiny cell) against an explicit image enumeration in numpy.
The cases follow the reference's fused-contact test table (apps/CCOptimization/tests/test_fused_contact.cpp:
387-418): skewed cells, radii spread over 0.7-1.3 times the mean, coincident particles, tiny cells."""
import math

import numpy as np
import pytest

from multi_cases import cell, radii_of, species_setup
from oracle import pyoracle as po


def total(s, a, idx, ks, vs, radii, u0=1.0, e0=1.0, rep=True):
    rc = 2.01 * float(np.max(radii)) if len(radii) else 0.0
    return po.multi_energy_forces(s, a, idx, ks, vs, radii, rc, u0, e0, rep)


@pytest.mark.parametrize("d,N", [(2, 120), (3, 90)])
def test_one_species_equal_radii_is_single_species_oracle(d, N):
    a = cell(N, d)
    s = po.make_config(7, N, d)
    k, _ = po.get_ks(a, po.getK(0.3, 1.0, d))
    v = np.ones(len(k))
    r = 0.31
    E, F, ES = po.multi_energy_forces(s, a, [list(range(N))], [k], [v], np.full(N, r), 2.01 * r, 1.0, 1.0, True)
    E1, F1 = po.total_energy_forces(s, a, k, v, 0.0, 2 * r, 1.0)
    assert abs(E - E1) <= 1e-12 * max(1.0, abs(E1))
    assert np.max(np.abs(F - F1)) <= 1e-12 * max(1.0, np.max(np.abs(F1)))
    # the CC part alone (contact off) equals the species energy and the plain CC oracle
    E0, F0, _ = po.multi_energy_forces(s, a, [list(range(N))], [k], [v], np.full(N, r), 2.01 * r, 1.0, 1.0, False)
    Ecc, Fcc = po.total_energy_forces(s, a, k, v)
    assert E0 == ES[0] and abs(E0 - Ecc) <= 1e-12 * abs(Ecc)
    assert np.max(np.abs(F0 - Fcc)) <= 1e-12 * np.max(np.abs(Fcc))


@pytest.mark.parametrize("d,skew,sizes,phi", [(2, True, (30, 50), 0.3), (3, False, (20, 25, 15), 0.3)])
def test_forces_are_minus_energy_gradient(d, skew, sizes, phi):
    N = sum(sizes)
    a = cell(N, d, skew)
    s = po.make_config(31 + N + d, N, d)
    idx, ks, vs = species_setup(a, sizes, 0.3)
    radii = radii_of(N, d, phi, True)
    u0, e0 = 0.8, 1.7
    E, F, _ = total(s, a, idx, ks, vs, radii, u0, e0)
    b = po.reciprocal_basis(a)
    h = 1e-6
    worst = 0.0
    for i in range(0, N, 7):
        for c in range(d):
            # a Cartesian step along axis c in relative coordinates: ds = e_c . b / (2 pi)
            ds = b[:, c] / (2 * math.pi) * h
            sp = s.copy()
            sm = s.copy()
            sp[i] += ds
            sm[i] -= ds
            Ep = total(sp, a, idx, ks, vs, radii, u0, e0)[0]
            Em = total(sm, a, idx, ks, vs, radii, u0, e0)[0]
            fd = -(Ep - Em) / (2 * h)
            worst = max(worst, abs(fd - F[i, c]))
    assert worst <= 1e-6 * max(1.0, np.max(np.abs(F)))


def test_momentum_and_relabelling():
    d, sizes = 2, (40, 35, 25)
    N = sum(sizes)
    a = cell(N, d, True)
    s = po.make_config(5, N, d)
    s[N - 1] = s[0]  # a coincident pair
    idx, ks, vs = species_setup(a, sizes, 0.35)
    radii = radii_of(N, d, 0.4, True)
    E, F, ES = total(s, a, idx, ks, vs, radii)
    assert np.max(np.abs(F.sum(axis=0))) <= 1e-9 * max(1.0, np.max(np.abs(F)))
    # relabel: particle i becomes perm position; species lists keep their order
    rng = np.random.default_rng(3)
    perm = rng.permutation(N)        # new index of old particle i is inv[i]
    inv = np.argsort(perm)
    s2 = s[perm]
    r2 = radii[perm]
    idx2 = [[int(inv[i]) for i in lst] for lst in idx]
    E2, F2, ES2 = total(s2, a, idx2, ks, vs, r2)
    assert np.array_equal(ES, ES2)  # species sums run in each species' own order: unchanged bits
    assert abs(E - E2) <= 1e-12 * max(1.0, abs(E))
    assert np.max(np.abs(F2 - F[perm])) <= 1e-12 * max(1.0, np.max(np.abs(F)))


def test_two_particles_closed_form():
    d = 3
    a = np.eye(d) * 40.0
    r1, r2 = 0.7, 1.1
    sig = r1 + r2
    x = 1.3  # centre distance, below sigma
    s = np.array([[0.5, 0.5, 0.5], [0.5 + x / 40.0, 0.5, 0.5]])
    e0 = 2.5
    # no constraints: only the contact term
    E, F, _ = po.multi_energy_forces(s, a, [[0, 1]], [np.zeros((0, d))], [np.zeros(0)], np.array([r1, r2]),
                                     2.01 * r2, 1.0, e0, True)
    assert E == pytest.approx(e0 * (1 - x / sig) ** 2, rel=1e-14)
    fmag = e0 * 2.0 / sig * (1.0 / x - 1.0 / sig) * x  # along -x on particle 0
    assert F[0, 0] == pytest.approx(-fmag, rel=1e-13) and F[1, 0] == pytest.approx(fmag, rel=1e-13)
    assert abs(F[0, 1]) + abs(F[0, 2]) < 1e-15
    # beyond contact: nothing
    s[1, 0] = 0.5 + (sig * 1.0001) / 40.0
    E, F, _ = po.multi_energy_forces(s, a, [[0, 1]], [np.zeros((0, d))], [np.zeros(0)], np.array([r1, r2]),
                                     2.01 * r2, 1.0, e0, True)
    assert E == 0.0 and not F.any()


def test_tiny_cell_contacts_through_several_images():
    d, N = 2, 7
    a = cell(N, d, True)
    s = po.make_config(44, N, d)
    radii = radii_of(N, d, 3.0, True)  # dense: pairs touch through two or more images
    rc = 2.01 * radii.max()
    slots, F = po.contact_radii(s, a, radii, rc, 1.0)
    # explicit enumeration over a generous image range
    r = s @ a
    E_ref = 0.0
    F_ref = np.zeros((N, d))
    multi = 0
    for i in range(N):
        for j in range(N):
            if i == j:
                continue
            sig = radii[i] + radii[j]
            hits = 0
            for n0 in range(-4, 5):
                for n1 in range(-4, 5):
                    dv = r[j] - r[i] + n0 * a[0] + n1 * a[1]
                    rr = float(np.sqrt(dv @ dv))
                    if rr < sig:
                        hits += 1
                        if i < j:
                            E_ref += (1 - rr / sig) ** 2
                        F_ref[i] += 2.0 / sig * (1.0 / rr - 1.0 / sig) * (-dv)
            multi = max(multi, hits)
    assert multi >= 2
    assert slots.sum() == pytest.approx(E_ref, rel=1e-12)
    assert np.max(np.abs(F - F_ref)) <= 1e-12 * np.max(np.abs(F_ref))
