"""Host mirror of the reference's Potential interface for the collective-coordinate path.

Names, argument meaning and error behaviour follow jaeukk/modules:
  * Configuration / GetUnitCubicBox            core/PeriodicCellList.h:46, core/PeriodicCellList.cpp:106-118
  * GetKs, getK, get_chi                        core/PeriodicCellList.cpp:509-735, apps/CCOptimization/main.cpp:2018-2026
  * ShiftedCCPotential                          potential/CollectiveCoordinatePotential.h:53-1497
  * RepulsiveCCPotential (soft core + CC)       apps/CCOptimization/RepulsiveCCPotential.h:21-157
Every evaluation runs in the native engine (modules_amd/libccgpu.so) on the GPU.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import check, dp, ip32, lib

pi = 3.1415926535897932384626433832795028841971693993751058209  # core/etc.h:9
Potential_Check_Delta = 1e-7  # potential/Potential.cpp:5


def _p(a):
    return a.ctypes.data_as(dp)


def _precision_code(mode):
    codes = {"fp64": 0, "mixed": 1, 0: 0, 1: 1}
    if mode not in codes:
        raise ValueError('precision must be "fp64" or "mixed"')
    return codes[mode]


class Configuration:
    """Periodic point pattern: basis vectors (rows) and relative coordinates in [0,1)."""

    def __init__(self, basis, relative=None):
        self.basis = np.array(basis, dtype=np.float64, copy=True)
        d = self.basis.shape[0]
        if self.basis.shape != (d, d) or not 1 <= d <= 4:
            raise ValueError("basis must be d x d with 1 <= d <= 4")
        self.relative = np.zeros((0, d)) if relative is None else np.array(relative, dtype=np.float64).reshape(-1, d)
        self.relative = _wrap(self.relative)

    def GetDimension(self):
        return self.basis.shape[0]

    def NumParticle(self):
        return self.relative.shape[0]

    def GetBasisVector(self, i):
        return self.basis[i].copy()

    def GetRelativeCoordinates(self, i):
        return self.relative[i].copy()

    def GetCartesianCoordinates(self, i):
        return self.relative[i] @ self.basis

    def PeriodicVolume(self):
        return lib().ccgpu_cell_volume(self.GetDimension(), _p(np.ascontiguousarray(self.basis)))

    def GetReciprocalBasisVector(self, i):
        """GetReciprocalBasisVector (core/PeriodicCellList.h:999-1031): b_i with a_j . b_i = 2 pi delta_ij."""
        d = self.GetDimension()
        b = np.zeros((d, d))
        check(lib().ccgpu_reciprocal_basis(d, _p(np.ascontiguousarray(self.basis)), _p(b)))
        return b[i]

    def CartesianCoord2RelativeCoord(self, r):
        """core/PeriodicCellList.h:1223-1230: s_i = (r . b_i) * (0.5 / pi)."""
        r = np.asarray(r, dtype=np.float64)
        return np.array([np.dot(r, self.GetReciprocalBasisVector(i)) for i in range(self.GetDimension())]) * (0.5 / pi)

    def Insert(self, relative):
        """Insert (core/PeriodicCellList.h:517-538): wrap the relative coordinate into [0,1) twice."""
        r = _wrap(np.asarray(relative, dtype=np.float64).reshape(1, -1))
        self.relative = np.vstack([self.relative, r])

    def MoveParticle(self, i, relative):
        self.relative[i] = _wrap(np.asarray(relative, dtype=np.float64).reshape(1, -1))[0]

    def Rescale(self, factor):
        self.basis = self.basis * factor

    def copy(self):
        c = Configuration(self.basis, self.relative)
        for attr in ("cell_size", "names"):
            if hasattr(self, attr):
                setattr(c, attr, getattr(self, attr))
        return c


def _wrap(s):
    y = s - np.floor(s)
    return y - np.floor(y)


def GetUnitCubicBox(d, CellSize=0.1):
    return Configuration(np.eye(d))


class RandomGenerator:
    """RandomGenerator(seed) (core/RandomGenerator.h:10-30): MT19937 with RandomDouble() = u32 / 2^32
    (core/RandomGenerator.cpp:24-27).  The draws come from the engine's MT19937 (ccgpu_random_configuration
    with d = 1); the generator keeps its position, so successive calls continue one stream."""

    def __init__(self, seed=0):
        self.seed = int(seed) & 0xFFFFFFFF
        self._buf = np.zeros(0)
        self._pos = 0

    def _ensure(self, n):
        need = self._pos + n
        if need > self._buf.size:
            size = max(need, 2 * self._buf.size, 4096)
            buf = np.zeros(size)
            check(lib().ccgpu_random_configuration(self.seed, size, 1, _p(buf)))
            self._buf = buf

    def RandomDouble(self):
        self._ensure(1)
        v = float(self._buf[self._pos])
        self._pos += 1
        return v

    def RandomDoubles(self, n):
        """The next n RandomDouble() values, in draw order."""
        self._ensure(n)
        v = self._buf[self._pos:self._pos + n].copy()
        self._pos += n
        return v


def RandomConfiguration(seed, N, d, L=None):
    """GetUnitCubicBox(d) rescaled by L (default pow(N, 1.0/d), density 1) with N particles drawn from
    RandomGenerator(seed) (apps/CCOptimization/main.cpp:198, 207-215).  `seed` may also be a RandomGenerator,
    whose stream then continues (particle i takes its next draws i*d .. i*d+d-1)."""
    if isinstance(seed, RandomGenerator):
        s = seed.RandomDoubles(N * d).reshape(N, d)
    else:
        s = np.zeros((N, d))
        check(lib().ccgpu_random_configuration(seed & 0xFFFFFFFF, N, d, _p(s)))
    if L is None:
        L = math.pow(N, 1.0 / d)
    c = Configuration(np.eye(d) * L, s)
    c.cell_size = 0.1 * L  # GetUnitCubicBox(d, 0.1) rescaled by L (core/PeriodicCellList.h:1201-1210)
    return c


def getK(chi, rho, d):
    return lib().ccgpu_get_k(chi, rho, d)


def get_chi(K, rho, d):
    c = {1: 2.0, 2: pi, 3: 4.1887902047863909846, 4: 4.9348022005446793094}[d]
    return c * K ** d / (2 * d * (2 * pi) ** d) / rho


def GetKs(config, CircularKMax, LinearKMax=None, SampleProbability=1.0, return_indices=False):
    """GetKs(tempList, CircularKMax, LinearKMax, SampleProbability): the wavevectors with |k| < K (half space).
    Only the LinearKMax == CircularKMax, SampleProbability >= 1 form used on the stealthy path is offered."""
    if LinearKMax is None:
        LinearKMax = CircularKMax
    if LinearKMax != CircularKMax or SampleProbability < 1.0:
        raise NotImplementedError("GetKs: only LinearKMax == CircularKMax and SampleProbability >= 1")
    basis = np.ascontiguousarray(config.basis if isinstance(config, Configuration) else config, dtype=np.float64)
    d = basis.shape[0]
    M = C.c_int64(0)
    check(lib().ccgpu_get_ks(d, _p(basis), CircularKMax, 0, None, None, C.byref(M)))
    k = np.zeros((max(M.value, 1), d))
    n = np.zeros((max(M.value, 1), d), dtype=np.int32)
    check(lib().ccgpu_get_ks(d, _p(basis), CircularKMax, M.value, _p(k), n.ctypes.data_as(ip32), C.byref(M)))
    k, n = k[: M.value], n[: M.value]
    return (k, n) if return_indices else k


class KPointConstraint:
    """One constraint: absolute wavevector k and weight V (CollectiveCoordinatePotential.h:84-94)."""

    __slots__ = ("k", "V")

    def __init__(self, k, V):
        self.k = np.asarray(k, dtype=np.float64).copy()
        self.V = float(V)


class ShiftedCCPotential:
    """Phi = (1/V) sum_k V_k |rho(k)|^2 + Shift on the GPU (CollectiveCoordinatePotential.h:53-1497)."""

    _S2 = False  # target-S(k) form (ShiftedCCPotential_S2)

    def __init__(self, dimension):
        self.Dimension = int(dimension)
        self.constraints = []
        self.S0 = []
        self.Shift = 0.0
        self.ParallelNumThread = 1  # kept for interface parity; the GPU does not use it
        self.RhoReal = np.zeros(0)
        self.RhoImag = np.zeros(0)
        self.RhoValid = False
        self._h = C.c_void_p()
        self._create()
        self._pushed_k = None
        self._pushed_v = None
        self._pushed_shift = None
        self._softcore = (0.0, 0.0)
        self._pushed_softcore = None
        self._config = None

    def _create(self):
        create = lib().ccgpu_potential_create_s2 if self._S2 else lib().ccgpu_potential_create
        check(create(self.Dimension, C.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            lib().ccgpu_potential_destroy(h)
            self._h = None

    # --- constraints -------------------------------------------------------------------------
    def AddConstraint(self, k, vals):
        """AddConstraint(k, vals): vals[0] is the weight; empty vals raises ValueError (std::invalid_argument).
        The target-S(k) form takes vals = [V(k), S0(k)] (CollectiveCoordinatePotential.h:1663-1668)."""
        if self._S2:
            if len(vals) < 2:
                raise ValueError("ShiftedCCPotential_S2::AddConstraint: vals must contain the weight V(k) and the "
                                 "target S0(k).")
            self.S0.append(float(vals[1]))
        elif len(vals) == 0:
            raise ValueError("ShiftedCCPotential::AddConstraint: vals must contain at least the weight V(k).")
        self.constraints.append(KPointConstraint(k, vals[0]))
        self.RhoValid = False

    def PrintConstraints(self, ofile=None):
        """PrintConstraints (CollectiveCoordinatePotential.h:1340-1345, 1672-1677)."""
        import sys
        out = ofile or sys.stdout
        for i, c in enumerate(self.constraints):
            ks = " ".join("%g" % x for x in c.k[: self.Dimension])
            if self._S2:
                out.write("@ k= %s: v=%g, S0(k)=%g\n" % (ks, c.V, self.S0[i]))
            else:
                out.write("@ k= %s: v=%g\n" % (ks, c.V))

    def _sync(self):
        d = self.Dimension
        M = len(self.constraints)
        k = np.zeros((M, d))
        v = np.zeros(M)
        for i, c in enumerate(self.constraints):
            k[i] = c.k[:d]
            v[i] = c.V
        if self._S2:
            if len(self.S0) != M:
                raise ValueError("ShiftedCCPotential_S2: one target S0 per constraint")
            v = np.concatenate([v, np.asarray(self.S0, dtype=np.float64)])  # weights then targets
        if self._pushed_k is None or not (np.array_equal(k, self._pushed_k) and np.array_equal(v, self._pushed_v)):
            if self._S2:
                check(lib().ccgpu_potential_set_constraints_s2(self._h, M, _p(np.ascontiguousarray(k)),
                                                               _p(np.ascontiguousarray(v[:M])),
                                                               _p(np.ascontiguousarray(v[M:]))))
            else:
                check(lib().ccgpu_potential_set_constraints(self._h, M, _p(np.ascontiguousarray(k)), _p(v)))
            self._pushed_k, self._pushed_v = k, v
            self.RhoValid = False
        if self._pushed_shift != self.Shift:
            check(lib().ccgpu_potential_set_shift(self._h, float(self.Shift)))
            self._pushed_shift = self.Shift
        if self._pushed_softcore != self._softcore:
            check(lib().ccgpu_potential_set_softcore(self._h, *self._softcore))
            self._pushed_softcore = self._softcore
            self.RhoValid = False

    # --- binding -----------------------------------------------------------------------------
    def SetConfiguration(self, config):
        """Bind a configuration and invalidate rho (CollectiveCoordinatePotential.h:1457-1463)."""
        if config.GetDimension() != self.Dimension:
            raise ValueError("dimension mismatch")
        self._config = config
        self._sync()
        rel = np.ascontiguousarray(config.relative, dtype=np.float64)
        check(lib().ccgpu_potential_set_configuration(self._h, _p(np.ascontiguousarray(config.basis)),
                                                      rel.shape[0], _p(rel)))
        self.RhoValid = False

    def SupportsIndexedView(self):
        """True only for the plain form (CollectiveCoordinatePotential.h:1465-1475)."""
        return type(self) in (ShiftedCCPotential, RepulsiveCCPotential) and not self._S2 and \
            getattr(self, "_softcore", (0.0, 0.0)) == (0.0, 0.0)

    def SetIndexedConfiguration(self, full, indices):
        """Bind the sub-configuration whose particle q is particle indices[q] of `full` (:1477-1496): the same bits
        as SetConfiguration on that sub-configuration.  RuntimeError (std::logic_error) for the derived forms,
        IndexError (std::out_of_range) for an index outside full."""
        if not self.SupportsIndexedView():
            raise RuntimeError("ShiftedCCPotential::SetIndexedConfiguration: not supported by this derived class; "
                               "use SetConfiguration.")
        idx = np.ascontiguousarray(indices, dtype=np.int64)
        if idx.size and (idx.min() < 0 or idx.max() >= full.NumParticle()):
            raise IndexError("ShiftedCCPotential::SetIndexedConfiguration: particle index out of range.")
        self._sync()
        rel = np.ascontiguousarray(full.relative, dtype=np.float64)
        check(lib().ccgpu_potential_set_indexed_configuration(
            self._h, _p(np.ascontiguousarray(full.basis)), rel.shape[0], _p(rel), idx.size,
            idx.ctypes.data_as(C.POINTER(C.c_int64))))
        self._config = Configuration(full.basis, rel[idx] if idx.size else np.zeros((0, full.GetDimension())))
        self.RhoValid = False

    def _require(self):
        if self._config is None:
            raise RuntimeError("no configuration bound: call SetConfiguration first")
        self._sync()

    def GetRho(self):
        self._require()
        M = len(self.constraints)
        re = np.zeros(max(M, 1))
        im = np.zeros(max(M, 1))
        check(lib().ccgpu_potential_get_rho(self._h, _p(re), _p(im)))
        self.RhoReal, self.RhoImag = re[:M], im[:M]
        self.RhoValid = True

    def Energy(self):
        self._require()
        E = C.c_double(0)
        check(lib().ccgpu_potential_energy(self._h, C.byref(E)))
        return E.value

    def Force(self, i):
        self._require()
        F = np.zeros(self.Dimension)
        check(lib().ccgpu_potential_force(self._h, int(i), _p(F)))
        return F

    def AllForce(self):
        self._require()
        F = np.zeros((self._config.NumParticle(), self.Dimension))
        check(lib().ccgpu_potential_all_force(self._h, _p(F)))
        return F

    def EnergyAndAllForce(self):
        """CombinedEnergyForce::EnergyAndAllForce (potential/CombinedEnergyForce.h:17-28)."""
        self._require()
        E = C.c_double(0)
        F = np.zeros((self._config.NumParticle(), self.Dimension))
        check(lib().ccgpu_potential_energy_and_all_force(self._h, C.byref(E), _p(F)))
        return E.value, F

    def Check(self, c, Nbr):
        """Potential::Check (potential/Potential.cpp:5-19): one-sided finite difference of the energy along x
        for particle Nbr, delta 1e-7; prints (and returns) (E0 - E1) / delta / F_x, close to 1 when Force is
        consistent with Energy and c is not a minimum."""
        c = c.copy()
        self.SetConfiguration(c)
        force = self.Force(Nbr)
        E0 = self.Energy()
        temp = c.GetCartesianCoordinates(Nbr).copy()
        temp[0] += Potential_Check_Delta
        c.MoveParticle(Nbr, c.CartesianCoord2RelativeCoord(temp))
        self.SetConfiguration(c)
        E1 = self.Energy()
        ratio = (E0 - E1) / Potential_Check_Delta / force[0]
        print(ratio)
        return ratio

    def SetTiles(self, segment_len=0, rho_chunk=0, force_chunk=0):
        check(lib().ccgpu_potential_set_tiles(self._h, segment_len, rho_chunk, force_chunk))

    def SetRhoKernel(self, variant):
        check(lib().ccgpu_potential_set_rho_kernel(self._h, int(variant)))

    def SetPrecision(self, mode):
        """"fp64" (default) or "mixed": lattice phases and force terms in FP32 after an FP64 range reduction,
        sums in FP64 (no counterpart in the reference; its error against FP64 is the caller's to report)."""
        check(lib().ccgpu_potential_set_precision(self._h, _precision_code(mode)))
        self.RhoValid = False

    def SetRhoBackend(self, backend):
        """SetRhoBackend (CollectiveCoordinatePotential.h:162-169): "direct" (default); "nufft" -- the Type 1 /
        Type 2 NUFFT on a fine grid (the reference's FinufftType1 with Type 2 forces) when the constraint set is
        eligible and the cost gate (the autotuner's timed choice for the shape class, else the cost model; the
        reference's FinufftForceGate, :318-420) predicts a win, the direct sums otherwise; "nufft_always" -- the
        NUFFT whenever eligible.  GetLastRhoBackendUsed / GetLastFinufftStatus / NufftGateReason tell which and
        why."""
        codes = {"direct": 0, "nufft": 1, "finufft": 1, "FinufftType1": 1, "nufft_always": 2, 0: 0, 1: 1, 2: 2}
        if backend not in codes:
            raise ValueError('ShiftedCCPotential::SetRhoBackend: backend must be "direct", "nufft" or '
                             '"nufft_always"')
        self._rho_backend = codes[backend]
        self._push_backend()

    def SetFinufftTolerance(self, eps):
        """SetFinufftTolerance (:180-186): requested NUFFT tolerance, positive and finite (default 1e-12)."""
        if not (eps > 0.0) or not math.isfinite(eps):
            raise ValueError("ShiftedCCPotential::SetFinufftTolerance: eps must be positive and finite.")
        self._nufft_eps = float(eps)
        self._push_backend()

    def SetFinufftMaxModes(self, maxModes):
        """SetFinufftMaxModes (:199-206): cap on the mode box prod(2 K_c + 1) (default 2^22)."""
        if int(maxModes) < 1:
            raise ValueError("ShiftedCCPotential::SetFinufftMaxModes: maxModes must be >= 1.")
        self._nufft_max_modes = int(maxModes)
        self._push_backend()

    def _push_backend(self):
        check(lib().ccgpu_potential_set_rho_backend(self._h, getattr(self, "_rho_backend", 0),
                                                    getattr(self, "_nufft_eps", 1e-12),
                                                    getattr(self, "_nufft_max_modes", 1 << 22)))
        self.RhoValid = False

    def GetLastRhoBackendUsed(self):
        """"nufft" if the bound system evaluates through the NUFFT, else "direct"."""
        self._require()
        v = C.c_int(0)
        check(lib().ccgpu_potential_rho_backend_used(self._h, C.byref(v)))
        return "nufft" if v.value else "direct"

    def GetLastForceBackendUsed(self):
        """GetLastForceBackendUsed (CollectiveCoordinatePotential.h:275-285): "nufft" if the last all-force pass
        used the Type 2 transform, else "direct" (also before any)."""
        v = C.c_int(0)
        check(lib().ccgpu_potential_force_backend_used(self._h, C.byref(v)))
        return "nufft" if v.value else "direct"

    def _nufft_status(self):
        st, pd, pn = C.c_int(0), C.c_double(0), C.c_double(0)
        buf = C.create_string_buffer(256)
        check(lib().ccgpu_potential_nufft_status(self._h, C.byref(st), buf, 256, C.byref(pd), C.byref(pn)))
        return st.value, buf.value.decode(), pd.value, pn.value

    def GetLastFinufftStatus(self):
        """GetLastFinufftStatus (CollectiveCoordinatePotential.h:268-273): 0 when the last rho pass ran the NUFFT,
        -1 (FinufftNotCalled) otherwise."""
        return self._nufft_status()[0]

    def NufftGateReason(self):
        """Why the bound system takes the direct sums ("" when it takes the NUFFT) and the cost model's predicted
        microseconds of one E+F on each path: (reason, direct_us, nufft_us)."""
        self._require()
        _, why, pd, pn = self._nufft_status()
        return why, pd, pn

    def UsesLatticePath(self):
        v = C.c_int(0)
        check(lib().ccgpu_potential_uses_lattice_path(self._h, C.byref(v)))
        return bool(v.value)


class ShiftedCCPotential_S2(ShiftedCCPotential):
    """Phi = sum_k V_k (|rho(k)|^2 / N - S0_k)^2 + Shift on the GPU, forces
    -(4/N) sum_k V_k (S_k - S0_k) (Im rho cos(k.r) - Re rho sin(k.r)) k
    (ShiftedCCPotential_S2, CollectiveCoordinatePotential.h:1642-1748).  Constraints take vals = [V(k), S0(k)];
    the targets are in .S0, one per constraint."""

    _S2 = True

    @property
    def N(self):
        return 0 if self._config is None else self._config.NumParticle()


class ShiftedCCPotential_S2_MultiConfig(ShiftedCCPotential):
    """The configuration as NSubConfig sub-configurations of Np = N / NSubConfig consecutive particles, with
    Phi = sum_k V (<S(k)> - S0)^2 + Shift over the mean S(k) (ShiftedCCPotential_S2_MultiConfig,
    CollectiveCoordinatePotential.h:1756-1901).  Constraints take vals = [V(k), S0(k)]; GetRho fills RhoReal / RhoImag
    with NSubConfig x M values (sub-configuration c at c M) and Snow with the mean S(k)."""

    _S2 = True

    def __init__(self, d, NSubConfig):
        self.NSubConfig = int(NSubConfig)
        super().__init__(d)
        self.Snow = np.zeros(0)

    def _create(self):
        check(lib().ccgpu_potential_create_s2_multiconfig(self.Dimension, self.NSubConfig, C.byref(self._h)))

    @property
    def Np(self):
        return 0 if self._config is None else self._config.NumParticle() // self.NSubConfig

    def GetRho(self):
        self._require()
        M = len(self.constraints)
        n = max(M * self.NSubConfig, 1)
        re = np.zeros(n)
        im = np.zeros(n)
        check(lib().ccgpu_potential_get_rho(self._h, _p(re), _p(im)))
        self.RhoReal, self.RhoImag = re[: M * self.NSubConfig], im[: M * self.NSubConfig]
        s2 = (self.RhoReal ** 2 + self.RhoImag ** 2).reshape(self.NSubConfig, M)
        self.Snow = s2.sum(axis=0) / (self.Np * self.NSubConfig) if M else np.zeros(0)
        self.RhoValid = True


def _copy_terms(dst, src):
    dst.constraints = [KPointConstraint(c.k, c.V) for c in src.constraints]
    dst.S0 = list(getattr(src, "S0", []))
    dst.Shift = src.Shift
    be = getattr(src, "_rho_backend", None)
    if be is not None:
        dst._rho_backend = be
        dst._nufft_eps = getattr(src, "_nufft_eps", 1e-12)
        dst._nufft_max_modes = getattr(src, "_nufft_max_modes", 1 << 22)
        dst._push_backend()


class ShiftedCCPotential_withWeight(ShiftedCCPotential):
    """rho(k) = sum_j w_j exp(i k.r_j), particle j's force times w_j (ShiftedCCPotential_withWeight,
    CollectiveCoordinatePotential.h:1498-1600): built from a plain potential and one weight per particle."""

    def __init__(self, Pot, weights):
        super().__init__(Pot.Dimension)
        _copy_terms(self, Pot)
        self.weights = np.ascontiguousarray(weights, dtype=np.float64)
        check(lib().ccgpu_potential_set_weights(self._h, len(self.weights), _p(self.weights)))


class LogCCPotential(ShiftedCCPotential):
    """Energy log(Phi) of the plain form, forces F / Phi (LogCCPotential, CollectiveCoordinatePotential.h:1604-1634):
    built from a plain potential."""

    def __init__(self, pot):
        super().__init__(pot.Dimension)
        _copy_terms(self, pot)
        check(lib().ccgpu_potential_set_log_form(self._h, 1))


class ShiftedCCPotential_ForMCMixin(ShiftedCCPotential):
    """A plain potential with Monte Carlo single-particle trial moves (ShiftedCCPotential_ForMCMixin,
    CollectiveCoordinatePotential.h:1906-1958): TryMove(prevCart, afterCart[, i]) returns the energy change from the
    incrementally updated rho (on the GPU), AcceptMove adopts it.  Passing the particle index i also moves it in the
    bound configuration."""

    def __init__(self, src):
        self._S2 = getattr(src, "_S2", False)
        super().__init__(src.Dimension)
        _copy_terms(self, src)

    def TryMove(self, prevCart, afterCart, i=-1):
        self._require()
        dE = C.c_double(0)
        pc = np.ascontiguousarray(prevCart, dtype=np.float64)[: self.Dimension]
        ac = np.ascontiguousarray(afterCart, dtype=np.float64)[: self.Dimension]
        check(lib().ccgpu_potential_try_move(self._h, int(i), _p(pc), _p(ac), C.byref(dE)))
        return dE.value

    def AcceptMove(self):
        """Adopt the trial rho (the caller moves the particle in its Configuration, as potential/Move.h:224-233)."""
        check(lib().ccgpu_potential_accept_move(self._h))
        self.RhoValid = False


class RepulsiveCCPotential(ShiftedCCPotential):
    """Soft core e0 (1 - r/sigma)^2 for r < sigma plus the CC potential
    (apps/CCOptimization/RepulsiveCCPotential.h:21-157).  Constraints go to .CCPotential, as in the reference."""

    def __init__(self, dim, epsilon, sigma, S0=0.0):
        # S0 selects the kernel (RepulsiveCCPotential.h:40-46): nonzero makes the CC part ShiftedCCPotential_S2,
        # whose constraints then take vals = [V(k), S0(k)]
        self._S2 = S0 != 0.0
        super().__init__(dim)
        self.e0 = float(epsilon)
        self.sigma = float(sigma)
        self.WithRepulsion = abs(self.e0) > 1e-10
        self._softcore = (self.e0, self.sigma) if self.WithRepulsion else (0.0, 0.0)
        self.CCPotential = self  # constraints live on the same engine object
