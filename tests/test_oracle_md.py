"""Pins of the MD oracle's thermostats (oracle/md.py) by what the mathematics fixes, on the CPU.

Nose-Hoover (potential/MD_System.cpp:314-368): the first half-kick carries no friction and the implicit second one
divides by 1 + xi dt / 2, so per step the friction is xi dt / 2 -- the dynamics dv/dt = a - (xi / 2) v with
dxi/dt = (sum m v^2 - N kT) / Q.  Along them d(Ek + U)/dt = -xi Ek, so
    H = Ek + U + Q xi^2 / 4 + (N kT / 2) eta,    deta/dt = xi,
is conserved, and the scheme keeps it to first order in dt.  A planted mistake (the target 2 kT written as kT)
breaks it, and so does the textbook extended energy of full-friction Nose-Hoover.

Andersen (MD_System.cpp:203-252): no collision is velocity Verlet itself; a collision draws the particle index and
each component by Box-Muller from the generator's stream, which a closed form checks on fixed draws; the redrawn
components follow N(0, kT / m).
"""
import numpy as np
import pytest

from oracle import md as omd
from oracle import pyoracle as po

N, D = 40, 2


@pytest.fixture(scope="module")
def system():
    rng = np.random.RandomState(3)
    L = np.sqrt(N)
    basis = np.eye(D) * L
    s = rng.rand(N, D)
    K = 2 * np.pi / L * np.sqrt(0.3 * D * (N - 1) / np.pi)  # about chi = 0.3
    k, _ = po.get_ks(basis, K)
    v = np.ones(len(k))

    def ef(x):
        return po.total_energy_forces(x, basis, k, v)

    return basis, s, ef


def nh_invariant(system, dt, T=4.0, kT=0.02, Q=1.0, md_cls=omd.MD, full_friction_form=False):
    basis, s, ef = system
    m = md_cls(basis, s, dt, 1.0, lambda x: ef(x)[1])
    m.Q = Q
    m.v = np.random.RandomState(1).normal(size=(N, D)) * 0.3
    eta, H, EU = 0.0, [], []
    for _ in range(int(round(T / dt))):
        xi0 = m.xi
        m.nose_hoover(1, kT)
        eta += 0.5 * (xi0 + m.xi) * dt
        U = ef(m.s)[0]
        Ek = m.kinetic()
        if full_friction_form:
            H.append(Ek + U + Q * m.xi ** 2 / 2 + N * kT * eta)
        else:
            H.append(Ek + U + Q * m.xi ** 2 / 4 + N * kT / 2 * eta)
        EU.append(Ek + U)
    H, EU = np.array(H), np.array(EU)
    return np.ptp(H), np.ptp(EU)


def test_nose_hoover_conserves_its_extended_energy_to_first_order(system):
    spread = {dt: nh_invariant(system, dt) for dt in (0.02, 0.01, 0.005)}
    print({dt: f"{h:.3e} (Ek+U {e:.3e})" for dt, (h, e) in spread.items()})
    for dt, (h, e) in spread.items():
        assert h < 0.03 * e  # the thermostat moves Ek + U by far more than H drifts
    # first order in dt: halving the step halves the drift
    assert 0.4 < spread[0.01][0] / spread[0.02][0] < 0.6
    assert 0.4 < spread[0.005][0] / spread[0.01][0] < 0.6


def test_nose_hoover_invariant_catches_planted_mistakes(system):
    good, moved = nh_invariant(system, 0.01)

    class WrongTarget(omd.MD):
        # the thermostat's target written as kT instead of 2 kT per particle (a plausible slip)
        def nose_hoover(self, repeat, kT):
            self._ensure()
            for _ in range(repeat):
                self.v = self.v + 0.5 * self.dt * self.acc
                half = self.v.copy()
                self._drift()
                self._get_acc()
                oldxi = self.xi
                for _ in range(20):
                    newv = (half + self.acc * 0.5 * self.dt) * (1.0 / (1 + self.xi * 0.5 * self.dt))
                    temp = oldxi
                    for i in range(self.N):
                        temp += (self.mass[i] * half[i] @ half[i] + self.mass[i] * newv[i] @ newv[i] - kT) / self.Q \
                            * 0.5 * self.dt
                    self.xi = temp
                self.v = newv
                self.time += self.dt

    bad, _ = nh_invariant(system, 0.01, md_cls=WrongTarget)
    assert bad > 20 * good
    # the full-friction extended energy is not this scheme's invariant
    textbook, _ = nh_invariant(system, 0.01, full_friction_form=True)
    assert textbook > 20 * good


def test_andersen_without_collisions_is_velocity_verlet(system):
    basis, s, ef = system
    a = omd.MD(basis, s, 0.01, 1.0, lambda x: ef(x)[1])
    b = omd.MD(basis, s, 0.01, 1.0, lambda x: ef(x)[1])
    v0 = np.random.RandomState(2).normal(size=(N, D)) * 0.2
    a.v, b.v = v0.copy(), v0.copy()
    used = a.anderson(25, 0.1, 0.0, np.zeros(4))
    b.evolve(25)
    assert used == 0 and np.array_equal(a.s, b.s) and np.array_equal(a.v, b.v)


def test_andersen_collision_draws_follow_box_muller(system):
    basis, s, ef = system
    m = omd.MD(basis, s, 0.01, 2.0, lambda x: np.zeros_like(x))  # free particles: only the collisions act
    kT = 0.3
    draws = np.array([0.51, 0.25, 0.125, 0.75, 0.6])  # particle floor(0.51 N); (t, u) per component
    used = m.anderson(1, kT, 1.0, draws)
    i = int(np.floor(0.51 * N))
    sigma = np.sqrt(kT / 2.0)
    want = [np.sqrt(-2 * np.log(0.25)) * np.sin(2 * np.pi * 0.125) * sigma,
            np.sqrt(-2 * np.log(0.75)) * np.sin(2 * np.pi * 0.6) * sigma]
    assert used == 1 + 2 * D
    assert np.allclose(m.v[i], want, rtol=0, atol=1e-15)
    assert np.count_nonzero(m.v) == D  # nobody else moved


def test_andersen_redrawn_components_are_maxwellian(system):
    basis, s, _ = system
    m = omd.MD(basis, s, 0.01, 1.5, lambda x: np.zeros_like(x))
    kT = 0.4
    rng = np.random.RandomState(7)
    samples = []
    for _ in range(3000):
        draws = rng.rand(1 + 2 * D)
        draws[1:] = np.maximum(draws[1:], 1e-300)
        m.v[:] = 0.0
        m.anderson(1, kT, 1.0, draws)
        samples.append(m.v[m.v != 0.0])
    x = np.concatenate(samples)
    var = kT / 1.5
    # mean 0 and variance kT / m within five standard errors of the sample
    assert abs(x.mean()) < 5 * np.sqrt(var / len(x))
    assert abs(x.var() / var - 1.0) < 5 * np.sqrt(2.0 / len(x))
