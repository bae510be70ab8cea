"""CPU oracle of ParticleMolecularDynamics (numpy over the oracle's forces).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product in modules_amd/ never imports it.

Restates potential/MD_System.cpp: the constructor (:17-29), _GetAcc / _EnsureInitialAcceleration (:66-87),
Evolve (:135-170), AndersonEvolve (:203-252), NoseHooverEvolve (:314-368), GetKineticEnergy (:370-377),
SetRandomSpeed (:379-394) and RemoveOverallMomentum (:395-407), with Cartesian positions moved as the reference
moves them (GetCartesianCoordinates + dt v, CartesianCoord2RelativeCoord, MoveParticle's wrap).
"""
from __future__ import annotations

import numpy as np

from . import pyoracle as po


def _wrap(s):
    s = s - np.floor(s)
    return s - np.floor(s)


class MD:
    def __init__(self, basis, relative, dt, mass, forces):
        """forces(s) -> F (N x d) of the potential at relative coordinates s."""
        self.a_basis = np.array(basis, dtype=np.float64)
        self.s = _wrap(np.array(relative, dtype=np.float64))
        self.N, self.d = self.s.shape
        self.dt = float(dt)
        self.mass = np.full(self.N, float(mass)) if np.ndim(mass) == 0 else np.array(mass, dtype=np.float64)
        self.v = np.zeros((self.N, self.d))
        self.acc = np.zeros((self.N, self.d))
        self.accel = False
        self.forces = forces
        self.xi = 0.0
        self.Q = 0.0
        self.time = 0.0
        self.b = po.reciprocal_basis(self.a_basis)

    def _get_acc(self):
        self.acc = self.forces(self.s) / self.mass[:, None]

    def _ensure(self):
        if not self.accel:
            self._get_acc()
            self.accel = True

    def _drift(self):
        cart = self.s @ self.a_basis + self.dt * self.v
        self.s = _wrap(cart @ self.b.T / (2 * np.pi))

    def evolve(self, repeat):
        self._ensure()
        for _ in range(repeat):
            self.v = self.v + 0.5 * self.dt * self.acc
            self._drift()
            self._get_acc()
            self.v = self.v + 0.5 * self.dt * self.acc
            self.time += self.dt

    def anderson(self, repeat, kT, prob, draws):
        self._ensure()
        k = 0
        count = 0.0
        for _ in range(repeat):
            self.v = self.v + 0.5 * self.dt * self.acc
            self._drift()
            self._get_acc()
            self.v = self.v + 0.5 * self.dt * self.acc
            self.time += self.dt
            count += prob
            while count > 0:
                i = int(np.floor(draws[k] * self.N))
                k += 1
                sigma = np.sqrt(kT / self.mass[i])
                for c in range(self.d):
                    while True:
                        t = draws[k]
                        k += 1
                        if t > 0:
                            break
                    self.v[i, c] = np.sqrt(-2 * np.log(t)) * np.sin(2 * np.pi * draws[k]) * sigma
                    k += 1
                count -= 1.0
        return k

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
                    temp += (self.mass[i] * half[i] @ half[i] + self.mass[i] * newv[i] @ newv[i] - 2 * kT) / self.Q \
                        * 0.5 * self.dt
                self.xi = temp
            self.v = newv
            self.time += self.dt

    def kinetic(self):
        return 0.5 * float(np.sum(np.sum(self.v * self.v, axis=1) * self.mass))

    def set_random_speed(self, kT, draws):
        k = 0
        for i in range(self.N):
            sigma = np.sqrt(kT / self.mass[i])
            for c in range(self.d):
                while True:
                    x = np.sqrt(-2 * np.log(draws[k])) * np.sin(2 * np.pi * draws[k + 1]) * sigma
                    k += 2
                    if abs(x) <= 1e30:
                        break
                self.v[i, c] = x
        return k

    def remove_overall_momentum(self):
        self.v = self.v - self.v.mean(axis=0)
