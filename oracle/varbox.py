"""CPU oracle of the variable-cell CC potentials (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product in modules_amd/ never imports it.

Restates, in the reference's loop order:
  - the weight and target catalog                 potential/CCPotential_Kernels.h:18-340
  - the k-point list of a cell                    potential/CCPotential_varBox.h:129-208 (points strictly inside
                                                  rK; +-k pairs kept once, first non-zero coefficient positive)
  - dBi / dki                                     potential/CCPotential_varBox.h:61-114 (one linear solve per
                                                  reciprocal vector and basis component)
  - Shift of the three forms                      potential/CCPotential_varBox.h:232-245, 117-124, 392-406
  - EnergyDerivativeToBasisVectors of the forms   potential/CCPotential_varBox.h:248-290, 333-378, 410-466
  - the fixed-cell energies they extend           CollectiveCoordinatePotential.h:1351-1372 (plain, via
                                                  pyoracle.energy) and :1686-1712 (S2, via pyoracle.s2_energy)
The k-point order here is |k| then the coefficients; the reference visits them in its cell list's order, so
tests compare sets (sorted) and order-free sums.
"""
from __future__ import annotations

import itertools

import numpy as np

from . import pyoracle as po

FORMS = ("stealthy", "S2", "S2_noV0")


def vk(kind, k, Rc=1.0):
    """tildeVk of K2 / KM2 / KM3 / KM4 / LemniscateVkTemplate (CCPotential_Kernels.h:25-30, 56-61, 86-91,
    116-121, 147-160)."""
    k = np.asarray(k, dtype=np.float64)
    k2 = float(k @ k)
    if kind == "Lemniscate":
        x2my2 = k[0] * k[0] - k[1] * k[1]
        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = k2 * k2 / (Rc * Rc) / x2my2 if x2my2 != 0 else np.inf
        if ratio < 1 and x2my2 > 0:
            return (ratio - 1) ** 4
        return 0.0
    kl = np.sqrt(k2)
    if kind == "K2":
        return (1.0 - kl) ** 2
    t = 1 / kl - 1
    return {"KM2": t ** 2, "KM3": t ** 3, "KM4": t ** 4}[kind]


def grad_vk(kind, k, Rc=1.0):
    """NablatildeVk (CCPotential_Kernels.h:35-41, 66-72, 96-102, 126-132, 166-180), as the reference writes it."""
    k = np.asarray(k, dtype=np.float64)
    k2 = float(k @ k)
    if kind == "Lemniscate":
        x2my2 = k[0] * k[0] - k[1] * k[1]
        ratio = k2 * k2 / (Rc * Rc) / x2my2 if x2my2 != 0 else np.inf
        if ratio < 1 and x2my2 > 0:
            temp = ratio - 1
            common = 2 * temp / (Rc * Rc) / x2my2 / x2my2
            return np.array([4 * k[0] * k2 * x2my2 - 2 * k[0] * k2 * k2,
                             4 * k[1] * k2 * x2my2 + 2 * k[1] * k2 * k2]) * (common * 2 * temp * temp)
        return np.zeros(2)
    kl = np.sqrt(k2)
    if kind == "K2":
        return ((-2.0) * (1.0 - kl) / kl) * k
    t = 1 / kl - 1
    if kind == "KM2":
        return (-2.0) * t / kl / kl / kl * k
    if kind == "KM3":
        return (-3.0) * t * t / kl / kl / kl * k
    return (-4.0) * t * t * t / kl / kl / kl * k


def s0k(kind, k, n=1.0):
    """S0k of S0K2 / S0Kn / S0Klc / S0KM2 / S0KMLog (CCPotential_Kernels.h:201-204, 230-233, 259-262, 285-290,
    314-318)."""
    k = np.asarray(k, dtype=np.float64)
    k2 = float(k @ k)
    if kind == "K2":
        return k2
    if kind == "Kn":
        return k2 ** (0.5 * n)
    if kind == "Klc":
        return 0.5 * (k2 ** (0.5 * n) + 1)
    if kind == "KM2":
        return (1 / np.sqrt(k2) - 1) ** 2
    if kind == "KMLog":
        return -0.5 * np.log(k2)
    raise ValueError(kind)


def grad_s0k(kind, k, n=1.0):
    """NablaS0k (CCPotential_Kernels.h:209-212, 238-241, 267-270, 295-301, 323-326)."""
    k = np.asarray(k, dtype=np.float64)
    k2 = float(k @ k)
    if kind == "K2":
        return 2.0 * k
    if kind == "Kn":
        return n * k2 ** (0.5 * n - 1) * k
    if kind == "Klc":
        return 0.5 * n * k2 ** (0.5 * n - 1) * k
    if kind == "KM2":
        kl = np.sqrt(k2)
        return (-2.0) * (1 / kl - 1) / kl / kl / kl * k
    if kind == "KMLog":
        return k * (-1.0 / k2)
    raise ValueError(kind)


def kpoints(a, rK):
    """The constrained k-points of cell a: (k M x d, coefficients M x d), every n with 0 < |sum n_i b_i| < rK and
    the first non-zero n_i positive, ordered by |k|^2 then n.  Enumerates a box of coefficients large enough for
    any cell (|n_i| <= rK |a_i| / 2 pi + 1)."""
    a = np.asarray(a, dtype=np.float64)
    d = a.shape[0]
    b = po.reciprocal_basis(a)
    lim = [int(np.floor(rK * np.linalg.norm(a[i]) / (2 * np.pi))) + 1 for i in range(d)]
    out = []
    for n in itertools.product(*[range(-L, L + 1) for L in lim]):
        nz = [x for x in n if x != 0]
        if not nz or nz[0] < 0:
            continue
        k = np.zeros(d)
        for i in range(d):
            k += n[i] * b[i]
        k2 = float(k @ k)
        if k2 < rK * rK:
            out.append((k2, tuple(n), k))
    out.sort(key=lambda t: (t[0], t[1]))
    if not out:
        return np.zeros((0, d)), np.zeros((0, d), dtype=np.int64)
    return np.array([t[2] for t in out]), np.array([t[1] for t in out], dtype=np.int64)


def shift(form, N, vol, v, s0=None):
    """The Shift each form sets at SetConfiguration."""
    if form == "stealthy":
        return N * (N - 1) / vol / 2.0 - N / vol * float(np.sum(v))
    if form == "S2":
        return 0.0
    sh = 0.0
    for i in range(len(v)):
        sh -= (s0[i] - 1) * (s0[i] - 1) * v[i]
    return sh


class Spec:
    def __init__(self, form="stealthy", vk="K2", s0=None, rK=1.0, n=1.0):
        self.form, self.vk, self.s0, self.rK, self.n = form, vk, s0, float(rK), float(n)


def terms(spec, a, N):
    """What SetConfiguration builds for cell a and N particles: k, coefficients, weights, targets, Shift."""
    k, nn = kpoints(a, spec.rK)
    v = np.array([vk(spec.vk, q, spec.rK) for q in k])
    t = np.array([s0k(spec.s0, q, spec.n) for q in k]) if spec.form != "stealthy" else None
    return k, nn, v, t, shift(spec.form, N, po.volume(a), v, t)


def energy(spec, a, s):
    """Energy() after SetConfiguration(a, s); returns (E, rho_re, rho_im, k, coefficients, v, s0)."""
    s = np.asarray(s, dtype=np.float64)
    N = s.shape[0]
    k, nn, v, t, sh = terms(spec, a, N)
    re, im = po.rho(s, a, k)
    if spec.form == "stealthy":
        E = po.energy(v, re, im, po.volume(a), sh)
    else:
        E = po.s2_energy(v, t, re, im, N, sh)
    return E, re, im, k, nn, v, t


def dBi(a, b, l, m):
    """d b_i / d a_lm: the solution of A x = -b_i[m] e_l for each reciprocal vector i (CCPotential_varBox.h:61-97)."""
    d = a.shape[0]
    out = []
    for i in range(d):
        rhs = np.zeros(d)
        rhs[l] = -b[i][m]
        out.append(np.linalg.solve(a, rhs))
    return out


def derivative(spec, a, s, Pressure=0.0):
    """EnergyDerivativeToBasisVectors(grad, Pressure) after SetConfiguration(a, s): grad[n d + m]."""
    a = np.asarray(a, dtype=np.float64)
    d = a.shape[0]
    N = np.asarray(s).shape[0]
    E, re, im, k, nn, v, t = energy(spec, a, s)
    b = po.reciprocal_basis(a)
    vol = po.volume(a)
    grad = np.zeros(d * d)
    for n in range(d):
        for m in range(d):
            db = dBi(a, b, n, m)
            temp = 0.0
            for i in range(len(k)):
                dk = np.zeros(d)
                for j in range(d):
                    dk += nn[i][j] * db[j]
                gv = grad_vk(spec.vk, k[i], spec.rK) @ dk
                r2 = re[i] * re[i] + im[i] * im[i]
                if spec.form == "stealthy":
                    temp += gv * (r2 - N)
                else:
                    ds = r2 / N - t[i]
                    gs = grad_s0k(spec.s0, k[i], spec.n) @ dk
                    if spec.form == "S2":
                        temp += gv * ds * ds
                        temp += v[i] * (-2.0) * ds * gs
                    else:
                        temp += gv * (ds * ds - (t[i] - 1.0) * (t[i] - 1.0)) + v[i] * (1.0 - t[i] - ds) * 2.0 * gs
            if spec.form == "stealthy":
                temp /= vol
            grad[n * d + m] = temp
    inv = np.linalg.inv(a)
    coef = (Pressure - E / vol) * vol if spec.form == "stealthy" else Pressure * vol
    for j in range(d):
        for q in range(d):
            grad[j * d + q] += coef * inv[q, j]
    return grad


def forces(spec, a, s):
    """AllForce after SetConfiguration(a, s) (the fixed-cell form's forces over the cell's constraints)."""
    s = np.asarray(s, dtype=np.float64)
    E, re, im, k, nn, v, t = energy(spec, a, s)
    if spec.form == "stealthy":
        return E, po.forces(s, a, k, v, re, im)
    return E, po.s2_forces(s, a, k, v, t, re, im)


def relax(spec, a, s, Pressure, Switch, maxstep=10000, Emin=-1e10, mf=40):
    """RelaxStructure_NLOPT[_Emin] (potential/StructureOptimization.cpp:881-931, 1046-1084) with Switch 0/1/2 and
    the reference's own LuksanPLIS (oracle/_ref): Objective (:447-578) = E + P V with the gradient of the cell
    parameters from EnergyDerivativeToBasisVectors and of the particles -F.a_j / rescale; GetConfig's
    validity bounds (:383-437, :11-14) give MaxEnergy.  Returns (H, a, s, evaluations)."""
    a = np.array(a, dtype=np.float64)
    s = np.array(s, dtype=np.float64)
    d = a.shape[0]
    N = s.shape[0]
    nb = d * d if Switch != 0 else 0
    total = 0
    for rescale in (10000.0, 1.0, 0.01):
        if energy(spec, a, s)[0] < Emin:
            break

        def split(x):
            cell = x[:nb].reshape(d, d) if Switch != 0 else a
            pos = x[nb:].reshape(N, d) / rescale if Switch != 1 else s
            return cell, pos

        def valid(cell):
            if Switch == 0:
                return True
            if np.any(np.sum(cell * cell, axis=1) > 4000000.0):
                return False
            vol = po.volume(cell)
            return 0.01 <= vol <= 10000000.0

        def fun(x):
            g = np.zeros(len(x))
            cell, pos = split(x)
            if not valid(cell) or po.volume(cell) < 1e-4 ** d:
                return 1e10, g
            E, F = forces(spec, cell, pos)
            if Switch != 1:
                g[nb:] = (-(F @ cell.T) / rescale).reshape(-1)
            if Switch != 0:
                g[:nb] = derivative(spec, cell, pos, Pressure)
            return E + Pressure * po.volume(cell), g

        x0 = np.concatenate([a.reshape(-1) if Switch != 0 else np.zeros(0),
                             (s * rescale).reshape(-1) if Switch != 1 else np.zeros(0)])
        x, minf, st, ev, trace = po.ref_plis(fun, x0, Emin, maxstep, mf)
        total += ev
        cell, pos = split(x)
        if valid(cell):
            a = np.array(cell)
            if Switch != 1:
                s = np.asarray(pos) - np.floor(pos)
        if Emin > -1e10 and energy(spec, a, s)[0] < Emin:
            break
    return energy(spec, a, s)[0] + Pressure * po.volume(a), a, s, total
