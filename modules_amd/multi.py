"""Host mirror of the reference's multi-species potential over the native engine (ccgpu_multi_* of include/ccgpu.h).

  * SpherePacking                  core/PeriodicCellList.h (Configuration with one radius per particle)
  * MultiStealthyPotential         apps/CCOptimization/MultiStealthyPotential.h:117-478,
                                   apps/CCOptimization/MultiStealthyPotential.cpp:139-985
Species potentials are views: ``species_constraints[s].idx`` is the species' particle list (assignment order) and
``species_constraints[s].CCPotential`` holds its constraints, densities and energy, as in the reference.  Every
evaluation runs on the GPU; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check, lib
from .potential import Configuration, KPointConstraint, _p, _wrap


class SpherePacking(Configuration):
    """Periodic packing: a Configuration whose particles carry a radius (the packing's characteristics)."""

    def __init__(self, basis, relative=None, radii=None):
        super().__init__(basis, relative)
        n = self.NumParticle()
        self.radii = np.zeros(n) if radii is None else np.array(radii, dtype=np.float64).reshape(n)

    def Insert(self, radius, relative):
        """SpherePacking::Insert(radius, relative): wrap into [0,1) and append."""
        super().Insert(relative)
        self.radii = np.append(self.radii, float(radius))

    def GetCharacteristics(self, i):
        return float(self.radii[i])

    def copy(self):
        return SpherePacking(self.basis, self.relative, self.radii)


class _SpeciesCC:
    """The species' ShiftedCCPotential as callers see it: constraints, RhoReal/RhoImag after GetRho, Energy()."""

    def __init__(self, owner, s):
        self._owner = owner
        self._s = s
        self.constraints = []
        self.RhoReal = np.zeros(0)
        self.RhoImag = np.zeros(0)
        self.RhoValid = False
        self._pushed = None

    def AddConstraint(self, k, vals):
        if len(vals) == 0:
            raise ValueError("ShiftedCCPotential::AddConstraint: vals must contain at least the weight V(k).")
        self.constraints.append(KPointConstraint(k, vals[0]))
        self.RhoValid = False

    def _sync(self):
        d = self._owner.Dimension
        M = len(self.constraints)
        k = np.zeros((M, d))
        v = np.zeros(M)
        for i, c in enumerate(self.constraints):
            k[i] = c.k[:d]
            v[i] = c.V
        if self._pushed is None or not (np.array_equal(k, self._pushed[0]) and np.array_equal(v, self._pushed[1])):
            check(lib().ccgpu_multi_set_species_constraints(self._owner._h, self._s, M, _p(np.ascontiguousarray(k)),
                                                            _p(v)))
            self._pushed = (k, v)
            self.RhoValid = False

    def GetRho(self):
        self._owner.GetRho(self._s)

    def Energy(self):
        self._owner._require()
        E = C.c_double(0)
        check(lib().ccgpu_multi_species_energy(self._owner._h, self._s, C.byref(E)))
        return E.value


class _Species:
    """species_potential (MultiStealthyPotential.h:68-110): the species' potential and its particle list."""

    def __init__(self, owner, s):
        self._owner = owner
        self._s = s
        self.CCPotential = _SpeciesCC(owner, s)

    @property
    def idx(self):
        n = C.c_int64(0)
        check(lib().ccgpu_multi_species_indices(self._owner._h, self._s, 0, None, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.int64)
        check(lib().ccgpu_multi_species_indices(self._owner._h, self._s, n.value,
                                                out.ctypes.data_as(_lib.i64p), C.byref(n)))
        return [int(i) for i in out[: n.value]]

    def PrintConstraints(self, ofile=None):
        import sys
        f = ofile or sys.stdout
        f.write("indices:" + "".join(f"\t{i}" for i in self.idx) + "\n")
        for c in self.CCPotential.constraints:
            f.write(f"@ k= {' '.join(repr(float(x)) for x in c.k)}: v={c.V!r}\n")
        f.write("\n\n")


class MultiStealthyPotential:
    """u0 sum_s (N_s/N) E_s + e0 sum_{i<j} (1 - r/(r_i+r_j))^2 on the GPU (MultiStealthyPotential.h:117-478)."""

    def __init__(self, dimension, epsilon):
        self.Dimension = int(dimension)
        self.ParallelNumThread = 1  # kept for interface parity; the GPU does not use it
        self.ReuseConfigurationStorage = True
        self.e0 = float(epsilon)
        self.u0 = 1.0
        self.WithRepulsion = abs(self.e0) > 1e-10
        self.species_constraints = []
        self._h = C.c_void_p()
        check(lib().ccgpu_multi_create(self.Dimension, self.e0, C.byref(self._h)))
        self._pushed_scales = None
        self._config = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            lib().ccgpu_multi_destroy(h)
            self._h = None

    # --- species and constraints ----------------------------------------------------------------
    def SetNumSpecies(self, num):
        """SetNumSpecies(num): append num empty species (MultiStealthyPotential.h:355-360)."""
        check(lib().ccgpu_multi_set_num_species(self._h, int(num)))
        base = len(self.species_constraints)
        for q in range(int(num)):
            self.species_constraints.append(_Species(self, base + q))

    def SetSpeciesOfParticle(self, prt_idx, species):
        """SetSpeciesOfParticle (MultiStealthyPotential.cpp:139-170)."""
        if self._config is None:
            raise RuntimeError("no configuration bound: call SetConfiguration first")
        check(lib().ccgpu_multi_set_species_of_particle(self._h, int(prt_idx), int(species)))

    def SpeciesOf(self):
        """idx2species[].species of every particle (-1 unassigned)."""
        n = 0 if self._config is None else self._config.NumParticle()
        out = np.zeros(max(n, 1), dtype=np.int32)
        check(lib().ccgpu_multi_species_of(self._h, out.ctypes.data_as(_lib.ip32)))
        return out[:n]

    def AddConstraint(self, species, k, V):
        """AddConstraint(species, k, V) (MultiStealthyPotential.cpp:176-180); V[0] is the weight."""
        if not 0 <= int(species) < len(self.species_constraints):
            raise IndexError("species index out of range")
        self.species_constraints[int(species)].CCPotential.AddConstraint(k, V)

    def SetRhoBackend(self, backend):
        """Species of a mixture evaluate with the direct sums (on the GPU); the NUFFT backend is selected per
        single-species potential (ShiftedCCPotential.SetRhoBackend)."""
        if str(backend).lower() not in ("direct", "0", "rhobackend.direct"):
            raise ValueError("MultiStealthyPotential::SetRhoBackend: only the direct backend is available")

    def PrintConstraints(self, ofile=None):
        import sys
        f = ofile or sys.stdout
        radii = self.GetRadii()
        for s, sp in enumerate(self.species_constraints):
            f.write(f"speciecs {s}\n\t (particle index, radius)")
            for i in sp.idx:
                f.write(f"\t({i},\t{radii[i]!r})\n")
            f.write("\n")
            for c in sp.CCPotential.constraints:
                f.write(f"@ k= {' '.join(repr(float(x)) for x in c.k)}: v={c.V!r}\n")
            f.write("----- \n\n")

    # --- configuration --------------------------------------------------------------------------
    def SetConfiguration(self, config):
        """SetConfiguration (MultiStealthyPotential.cpp:236-278)."""
        if config.GetDimension() != self.Dimension:
            raise ValueError("dimension mismatch")
        rel = np.ascontiguousarray(config.relative, dtype=np.float64)
        check(lib().ccgpu_multi_set_configuration(self._h, _p(np.ascontiguousarray(config.basis)), rel.shape[0],
                                                  _p(rel)))
        self._config = config

    def SetSecureRadii(self, radii):
        """SetSecureRadii (MultiStealthyPotential.cpp:448-458): finite positive radii; cutoff 2.01 max r."""
        r = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        if self._config is None:
            raise RuntimeError("no configuration bound: call SetConfiguration first")
        status = lib().ccgpu_multi_set_radii(self._h, r.shape[0], _p(r))
        if status == _lib.ERR_ARG:
            raise ValueError(lib().ccgpu_last_error().decode())
        check(status)

    def SetPacking(self, pck):
        """SetPacking (MultiStealthyPotential.cpp:461-489): positions and radii of a sphere packing."""
        if pck.GetDimension() != self.Dimension:
            raise ValueError("dimension mismatch")
        rel = np.ascontiguousarray(pck.relative, dtype=np.float64)
        radii = np.ascontiguousarray(pck.radii, dtype=np.float64)
        check(lib().ccgpu_multi_set_packing(self._h, _p(np.ascontiguousarray(pck.basis)), rel.shape[0], _p(rel),
                                            _p(radii)))
        self._config = pck

    def PrepareContactCells(self):
        """The contact cell list is rebuilt on the GPU at every evaluation; nothing to prepare."""

    def UpdateSubconfigs(self):
        """Species are views of the stored configuration; nothing to rebind."""

    def GetRc(self):
        rc = C.c_double(0)
        check(lib().ccgpu_multi_get_rc(self._h, C.byref(rc)))
        return rc.value

    def GetRadii(self):
        n = 0 if self._config is None else self._config.NumParticle()
        out = np.zeros(max(n, 1))
        check(lib().ccgpu_multi_get_radii(self._h, _p(out)))
        return out[:n]

    def ExportConfiguration(self, ofile, format="text"):
        """ExportConfiguration (MultiStealthyPotential.cpp:28-131): the stored configuration, species labels and radii
        as the text table ("text") or an MSPB v1 snapshot ("binary"), written to `ofile` (a path, or a file object
        opened in text / binary mode); returns what was written."""
        codes = {"text": 0, "binary": 1, "Text": 0, "Binary": 1}
        if format not in codes:
            raise ValueError("ExportConfiguration: unknown format.")
        size = C.c_int64(0)
        check(lib().ccgpu_multi_export(self._h, codes[format], None, 0, C.byref(size)))
        buf = C.create_string_buffer(max(size.value, 1))
        check(lib().ccgpu_multi_export(self._h, codes[format], buf, size.value, C.byref(size)))
        data = buf.raw[: size.value]
        out = data.decode() if codes[format] == 0 else data
        if isinstance(ofile, (str, bytes, os.PathLike)):
            with open(ofile, "w" if codes[format] == 0 else "wb") as f:
                f.write(out)
        elif ofile is not None:
            ofile.write(out)
        return out

    def GetPackingSnapshot(self):
        """GetPackingSnapshot (MultiStealthyPotential.cpp:15-26): cell, relative coordinates and radii."""
        self._require_bound()
        return SpherePacking(self._config.basis, self._config.relative, self.GetRadii())

    # --- evaluation -----------------------------------------------------------------------------
    def _require_bound(self):
        if self._config is None:
            raise RuntimeError("no configuration bound: call SetConfiguration first")

    def _require(self):
        self._require_bound()
        for sp in self.species_constraints:
            sp.CCPotential._sync()
        scales = (float(self.u0), float(self.e0))
        if self._pushed_scales != scales:
            check(lib().ccgpu_multi_set_scales(self._h, *scales))
            self._pushed_scales = scales

    def GetRho(self, species):
        """GetRho(species) (MultiStealthyPotential.cpp:184-190): fills the species' RhoReal / RhoImag."""
        self._require()
        cc = self.species_constraints[int(species)].CCPotential
        M = len(cc.constraints)
        re = np.zeros(max(M, 1))
        im = np.zeros(max(M, 1))
        check(lib().ccgpu_multi_get_rho(self._h, int(species), _p(re), _p(im)))
        cc.RhoReal, cc.RhoImag = re[:M], im[:M]
        cc.RhoValid = True

    def Energy(self):
        self._require()
        E = C.c_double(0)
        check(lib().ccgpu_multi_energy(self._h, C.byref(E)))
        return E.value

    def Force(self, i):
        self._require()
        F = np.zeros(self.Dimension)
        check(lib().ccgpu_multi_force(self._h, int(i), _p(F)))
        return F

    def AllForce(self):
        self._require()
        F = np.zeros((self._config.NumParticle(), self.Dimension))
        check(lib().ccgpu_multi_all_force(self._h, _p(F)))
        return F

    def EnergyAndAllForce(self):
        """EnergyAndAllForce (MultiStealthyPotential.cpp:932-984): Energy() then AllForce() in one pass."""
        self._require()
        E = C.c_double(0)
        F = np.zeros((self._config.NumParticle(), self.Dimension))
        check(lib().ccgpu_multi_energy_and_all_force(self._h, C.byref(E), _p(F)))
        return E.value, F

    def Relax(self, maxstep=10000, Emin=-1e10, history=40):
        """RelaxStructure_NLOPT[_Emin] of the bound configuration with this potential (the relax inside
        CollectiveCoordinateMultiRun for GetMultiCCO, apps/CCOptimization/main.cpp:1565-1590); the relaxed
        configuration stays bound.  Returns (energy, evaluations, NLopt code of the last pass)."""
        self._require()
        rel = np.ascontiguousarray(self._config.relative, dtype=np.float64).copy()
        E = C.c_double(0)
        ev = C.c_int(0)
        st = C.c_int(0)
        check(lib().ccgpu_multi_relax(self._h, _p(rel), int(maxstep), float(Emin), int(history), C.byref(E),
                                      C.byref(ev), C.byref(st)))
        self._config.relative = _wrap(rel)
        return E.value, ev.value, st.value
