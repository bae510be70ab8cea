"""Variable-cell collective-coordinate potentials (potential/CCPotential_varBox.h) and the named kernel
combinations of potential/CCPotential_Kernels.h, over the library's variable-cell form (include/ccgpu.h).

Each SetConfiguration with a new cell regenerates the constrained k-points (0 < |k| < rK, one of each +-k pair),
their weights tildeV(k), the targets S0(k) of the target-S(k) forms and the Shift; Energy / Force / GetRho run on
the GPU as for the fixed-cell potentials, and EnergyDerivativeToBasisVectors contracts the GPU's rho with the
k-point derivatives.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, ip32, lib
from .potential import KPointConstraint, ShiftedCCPotential, _p

# forms, weights and targets (cc_varbox.h)
FORM = {"stealthy": 0, "S2": 1, "S2_noV0": 2}
VK = {"K2": 0, "KM2": 1, "KM3": 2, "KM4": 3, "Lemniscate": 4}
S0K = {None: 0, "K2": 1, "Kn": 2, "Klc": 3, "KM2": 4, "KMLog": 5}

Potential_Check_Delta = 1e-7  # potential/Potential.h


def KPoints(basis, rK):
    """The constrained k-points a variable-cell potential generates for the cell `basis` (rows are basis vectors):
    (k M x d, integer coefficients M x d), 0 < |k| < rK, one of each +-k pair, ordered by |k| then coefficients."""
    basis = np.ascontiguousarray(basis, dtype=np.float64)
    d = basis.shape[0]
    M = C.c_int64(0)
    check(lib().ccgpu_varbox_kpoints(d, _p(basis), float(rK), 0, None, None, C.byref(M)))
    k = np.zeros((max(M.value, 1), d))
    n = np.zeros((max(M.value, 1), d), dtype=np.int32)
    check(lib().ccgpu_varbox_kpoints(d, _p(basis), float(rK), M.value, _p(k), n.ctypes.data_as(ip32), C.byref(M)))
    return k[: M.value], n[: M.value]


class VarBoxPotential(ShiftedCCPotential):
    """ShiftedCCPotential_varBoxable<Base> with the weight `vk` and target `s0` mixins
    (CCPotential_varBox.h:31-209, CCPotential_Kernels.h:18-340).  form: "stealthy" (ShiftedCCPotential_varBox),
    "S2" (ShiftedCCPotential_S2_varBox) or "S2_noV0" (ShiftedCCPotential_S2_varBox_noV0)."""

    def __init__(self, d, form="stealthy", vk="K2", s0=None, rK=1.0, n=1.0):
        self._spec = (FORM[form], VK[vk], S0K[s0], float(rK), float(n))
        self._S2 = form != "stealthy"
        self.rK = float(rK)
        super().__init__(d)

    def _create(self):
        check(lib().ccgpu_potential_create_varbox(self.Dimension, *self._spec, C.byref(self._h)))

    def AddConstraint(self, k, vals):
        raise RuntimeError("the constraints of a variable-cell potential follow its cell")

    def _sync(self):
        # constraints and Shift are the library's (regenerated at SetConfiguration); nothing to push
        pass

    def SetConfiguration(self, config):
        if config.GetDimension() != self.Dimension:
            raise ValueError("dimension mismatch")
        self._config = config
        rel = np.ascontiguousarray(config.relative, dtype=np.float64)
        check(lib().ccgpu_potential_set_configuration(self._h, _p(np.ascontiguousarray(config.basis)),
                                                      rel.shape[0], _p(rel)))
        self.RhoValid = False
        self._pull()

    def _pull(self):
        """Mirror the regenerated constraints (k, tildeV, S0) and their coefficients on the reciprocal basis."""
        d = self.Dimension
        M = C.c_int64(0)
        check(lib().ccgpu_potential_num_constraints(self._h, C.byref(M)))
        m = M.value
        k = np.zeros((max(m, 1), d))
        v = np.zeros(max(m, 1))
        s0 = np.zeros(max(m, 1))
        nn = np.zeros((max(m, 1), d), dtype=np.int32)
        check(lib().ccgpu_potential_get_constraints(self._h, _p(k), _p(v), _p(s0) if self._S2 else None,
                                                    nn.ctypes.data_as(ip32)))
        self.constraints = [KPointConstraint(k[i], v[i]) for i in range(m)]
        self.kCoeffs = nn[:m].copy()
        self.S0 = list(s0[:m]) if self._S2 else []

    def EnergyDerivativeToBasisVectors(self, Pressure=0.0):
        """grad[n d + m] = d(Phi + P V) / d(component m of basis vector n), as a flat array of d*d values."""
        self._require()
        g = np.zeros(self.Dimension * self.Dimension)
        check(lib().ccgpu_potential_energy_derivative_to_basis(self._h, float(Pressure), _p(g)))
        return g

    def Check2(self, c, Nbr, Pressure=0.0, delta=Potential_Check_Delta):
        """Potential::Check2 (potential/Potential.cpp:21-40): one-sided difference of E + P V along basis component
        Nbr over the analytic derivative; prints (and returns) the ratio, close to 1 when they agree."""
        d = self.Dimension
        Nbr = Nbr % (d * d)
        c = c.copy()
        self.SetConfiguration(c)
        der = self.EnergyDerivativeToBasisVectors(Pressure)
        E0 = self.Energy() + Pressure * c.PeriodicVolume()
        basis = c.basis.copy()
        basis[Nbr // d, Nbr % d] += delta
        c.basis = basis
        self.SetConfiguration(c)
        E1 = self.Energy() + Pressure * c.PeriodicVolume()
        ratio = (E1 - E0) / delta / der[Nbr]
        print(ratio)
        return ratio


class ShiftedCCPotential_varBox(VarBoxPotential):
    """(1/V) sum tildeV |rho|^2 + N (N - 1) / (2 V) - (N / V) sum tildeV (CCPotential_varBox.h:215-291)."""

    def __init__(self, d, rK=1.0, vk="K2"):
        super().__init__(d, "stealthy", vk, None, rK)


def _named(form, vk, s0, has_n=False, doc=""):
    if has_n:
        def init(self, d, n, rK=1.0):
            VarBoxPotential.__init__(self, d, form, vk, s0, rK, n)
    else:
        def init(self, d, rK=1.0):
            VarBoxPotential.__init__(self, d, form, vk, s0, rK)
    return type("_", (VarBoxPotential,), {"__init__": init, "__doc__": doc})


# the typedefs of CCPotential_Kernels.h:191-192, 342-358
K2Potential = _named("stealthy", "K2", None, doc="K2Template<ShiftedCCPotential_varBox>")
KM3Potential = _named("stealthy", "KM3", None, doc="KM3Template<ShiftedCCPotential_varBox>")
K2K2Potential = _named("S2", "K2", "K2", doc="S0K2Template<K2Template<ShiftedCCPotential_S2_varBox>>")
K2K2Potential_noV0 = _named("S2_noV0", "K2", "K2", doc="S0K2Template<K2Template<..._S2_varBox_noV0>>")
KM3K2Potential_noV0 = _named("S2_noV0", "KM3", "K2", doc="S0K2Template<KM3Template<..._S2_varBox_noV0>>")
KM2K2Potential = _named("S2", "KM2", "K2", doc="S0K2Template<KM2Template<ShiftedCCPotential_S2_varBox>>")
KM3K2Potential = _named("S2", "KM3", "K2", doc="S0K2Template<KM3Template<ShiftedCCPotential_S2_varBox>>")
KM4K2Potential = _named("S2", "KM4", "K2", doc="S0K2Template<KM4Template<ShiftedCCPotential_S2_varBox>>")
KM3KM2Potential = _named("S2", "KM3", "KM2", doc="S0KM2Template<KM3Template<ShiftedCCPotential_S2_varBox>>")
KM3KMLogPotential = _named("S2", "KM3", "KMLog", doc="S0KMLogTemplate<KM3Template<ShiftedCCPotential_S2_varBox>>")
KM2KnPotential = _named("S2", "KM2", "Kn", True, doc="S0KnTemplate<KM2Template<...>>(d, n, rK)")
KM3KnPotential = _named("S2", "KM3", "Kn", True, doc="S0KnTemplate<KM3Template<...>>(d, n, rK)")
KM4KnPotential = _named("S2", "KM4", "Kn", True, doc="S0KnTemplate<KM4Template<...>>(d, n, rK)")
KM2KlcPotential = _named("S2", "KM2", "Klc", True, doc="S0KlcTemplate<KM2Template<...>>(d, n, rK)")
KM3KlcPotential = _named("S2", "KM3", "Klc", True, doc="S0KlcTemplate<KM3Template<...>>(d, n, rK)")
KM4KlcPotential = _named("S2", "KM4", "Klc", True, doc="S0KlcTemplate<KM4Template<...>>(d, n, rK)")
for _name, _cls in list(globals().items()):
    if _name.endswith(("Potential", "Potential_noV0")) and isinstance(_cls, type) and _cls.__name__ == "_":
        _cls.__name__ = _cls.__qualname__ = _name


class LemniscateVkPotential(VarBoxPotential):
    """LemniscateVkTemplate<ShiftedCCPotential_varBox>(Rc) (CCPotential_Kernels.h:140-186): 2D, rK = Rc."""

    def __init__(self, Rc):
        super().__init__(2, "stealthy", "Lemniscate", None, Rc)
