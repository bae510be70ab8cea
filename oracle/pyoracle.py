"""ctypes wrapper of the CPU oracle (oracle/cc_oracle.c) and of the reference LuksanPLIS under oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry and by bench.py's CPU-baseline leg, as
the checker.  The product in modules_amd/ never imports this module.

Every function mirrors one reference routine; the C file's header lists the reference file:line each
one follows.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libccoracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_plis.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

OBJECTIVE = C.CFUNCTYPE(C.c_double, C.c_uint, _dp, _dp, C.c_void_p)

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        L.orc_mt_sequence.argtypes = [C.c_uint32, C.c_int64, C.POINTER(C.c_uint32)]
        L.orc_make_config.argtypes = [C.c_uint32, C.c_int64, C.c_int, _dp]
        L.orc_reciprocal_basis.argtypes = [C.c_int, _dp, _dp]
        L.orc_volume.argtypes = [C.c_int, _dp]
        L.orc_volume.restype = C.c_double
        L.orc_getK.argtypes = [C.c_double, C.c_double, C.c_int]
        L.orc_getK.restype = C.c_double
        L.orc_get_ks.argtypes = [C.c_int, _dp, C.c_double, C.c_int64, _dp, _ip]
        L.orc_get_ks.restype = C.c_int64
        L.orc_rho.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int64, _dp, _dp, _dp, C.c_int]
        L.orc_energy.argtypes = [C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double]
        L.orc_energy.restype = C.c_double
        L.orc_forces.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int64, _dp, _dp, _dp, _dp, C.c_double, _dp, C.c_int]
        L.orc_softcore_energy.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_double, C.c_double, C.c_int]
        L.orc_softcore_energy.restype = C.c_double
        L.orc_softcore_forces.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_double, C.c_double, _dp, C.c_int]
        L.orc_total_energy_forces.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int64, _dp, _dp, C.c_double,
                                              C.c_double, C.c_double, _dp, C.c_int]
        L.orc_total_energy_forces.restype = C.c_double
        L.orc_relax_emin.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int64, _dp, _dp, C.c_double, C.c_double,
                                     C.c_double, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_int,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_relax_emin.restype = C.c_double
        L.orc_max_threads.restype = C.c_int
        _i64p = C.POINTER(C.c_int64)
        L.orc_contact_radii.argtypes = [C.c_int, C.c_int64, _dp, _dp, _dp, C.c_double, C.c_double, _dp, _dp,
                                        C.c_int]
        L.orc_multi_energy_forces.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int, _i64p, _i64p, _i64p, _dp, _dp,
                                              _dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp, C.c_int]
        L.orc_multi_energy_forces.restype = C.c_double
        L.orc_s2_multiconfig.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int64, _dp, _dp, _dp, C.c_int64,
                                         C.c_double, _dp, _dp, _dp, C.c_int]
        L.orc_s2_multiconfig.restype = C.c_double
        L.orc_rho_weighted.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int64, _dp, _dp, _dp, _dp, C.c_int]
        L.orc_s2_energy.argtypes = [C.c_int64, _dp, _dp, _dp, _dp, C.c_int64, C.c_double]
        L.orc_s2_energy.restype = C.c_double
        L.orc_s2_forces.argtypes = [C.c_int, C.c_int64, _dp, _dp, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref():
    """The reference's LuksanPLIS (potential/LuksanPLIS.cpp) compiled under oracle/_ref."""
    global _ref
    if _ref is None:
        if not os.path.exists(_REF_PATH):
            raise RuntimeError(f"reference build missing: {_REF_PATH} (run `make -C oracle ref`)")
        R = C.CDLL(_REF_PATH)
        R.ref_luksan_plis.argtypes = [C.c_int, OBJECTIVE, C.c_void_p, _dp, _dp, C.c_double, C.c_int, C.c_int,
                                      C.POINTER(C.c_int), C.c_int]
        R.ref_luksan_plis.restype = C.c_int
        _ref = R
    return _ref


def _p(a):
    return a.ctypes.data_as(_dp)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def mt_sequence(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint32)
    lib().orc_mt_sequence(seed, count, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def make_config(seed: int, N: int, d: int) -> np.ndarray:
    s = np.zeros((N, d))
    lib().orc_make_config(seed, N, d, _p(s))
    return s


def cubic_basis(N: int, d: int, density: float = 1.0) -> np.ndarray:
    """Basis of GetUnitCubicBox(d) rescaled by L = pow(N, 1.0/d) (main.cpp:198, 207-209) at density 1."""
    L = pow(N / density, 1.0 / d) if density != 1.0 else pow(N, 1.0 / d)
    return np.eye(d) * L


def reciprocal_basis(a: np.ndarray) -> np.ndarray:
    a = _c(a)
    d = a.shape[0]
    b = np.zeros((d, d))
    lib().orc_reciprocal_basis(d, _p(a), _p(b))
    return b


def volume(a: np.ndarray) -> float:
    a = _c(a)
    return lib().orc_volume(a.shape[0], _p(a))


def getK(chi: float, rho: float, d: int) -> float:
    return lib().orc_getK(chi, rho, d)


def get_ks(a: np.ndarray, K: float):
    """GetKs(config, K, K, 1): returns (k[M,d], n[M,d] int32)."""
    a = _c(a)
    d = a.shape[0]
    M = lib().orc_get_ks(d, _p(a), K, 0, None, None)
    k = np.zeros((max(M, 1), d))
    n = np.zeros((max(M, 1), d), dtype=np.int32)
    lib().orc_get_ks(d, _p(a), K, M, _p(k), n.ctypes.data_as(_ip))
    return k[:M], n[:M]


def rho(s, a, k, nthreads=0):
    s, a, k = _c(s), _c(a), _c(k)
    N, d = s.shape
    M = k.shape[0]
    re = np.zeros(M)
    im = np.zeros(M)
    lib().orc_rho(d, N, _p(s), _p(a), M, _p(k.reshape(M, d) if M else np.zeros((1, d))), _p(re), _p(im),
                  nthreads or lib().orc_max_threads())
    return re, im


def energy(v, re, im, vol, shift=0.0):
    v, re, im = _c(v), _c(re), _c(im)
    return lib().orc_energy(len(v), _p(v), _p(re), _p(im), vol, shift)


def forces(s, a, k, v, re, im, nthreads=0):
    s, a, k, v, re, im = _c(s), _c(a), _c(k), _c(v), _c(re), _c(im)
    N, d = s.shape
    M = k.shape[0]
    F = np.zeros((N, d))
    if N == 0:
        return F
    kk = k if M else np.zeros((1, d))
    lib().orc_forces(d, N, _p(s), _p(a), M, _p(kk), _p(v if M else np.zeros(1)), _p(re if M else np.zeros(1)),
                     _p(im if M else np.zeros(1)), volume(a), _p(F), nthreads or lib().orc_max_threads())
    return F


def rho_weighted(s, a, k, w, nthreads=0):
    """ShiftedCCPotential_withWeight::GetRho: rho = sum_j w_j exp(i k.r_j)."""
    s, a, k, w = _c(s), _c(a), _c(k), _c(w)
    N, d = s.shape
    M = k.shape[0]
    re = np.zeros(M)
    im = np.zeros(M)
    if M:
        lib().orc_rho_weighted(d, N, _p(s), _p(a), M, _p(k), _p(w if N else np.zeros(1)), _p(re), _p(im),
                               nthreads or lib().orc_max_threads())
    return re, im


def weighted_forces(s, a, k, v, w, re, im, nthreads=0):
    """ShiftedCCPotential_withWeight::Force for every particle: the plain form's force over the weighted rho,
    times the particle's weight (CollectiveCoordinatePotential.h:1555-1577)."""
    return forces(s, a, k, v, re, im, nthreads) * _c(w)[:, None]


def s2_multiconfig(s, a, k, v, s0, nsub, shift=0.0, nthreads=0):
    """ShiftedCCPotential_S2_MultiConfig: (E, F, re nsub x M, im nsub x M)."""
    s, a, k, v, s0 = _c(s), _c(a), _c(k), _c(v), _c(s0)
    N, d = s.shape
    M = k.shape[0]
    re = np.zeros(max(nsub * M, 1))
    im = np.zeros(max(nsub * M, 1))
    F = np.zeros((max(N, 1), d))
    E = lib().orc_s2_multiconfig(d, N, _p(s), _p(a), M, _p(k if M else np.zeros((1, d))),
                                 _p(v if M else np.zeros(1)), _p(s0 if M else np.zeros(1)), int(nsub), shift,
                                 _p(re), _p(im), _p(F), nthreads or lib().orc_max_threads())
    return E, F[:N], re[: nsub * M].reshape(nsub, M), im[: nsub * M].reshape(nsub, M)


def s2_energy(v, s0, re, im, N, shift=0.0):
    """ShiftedCCPotential_S2::Energy: sum V (|rho|^2 / N - S0)^2 + Shift."""
    v, s0, re, im = _c(v), _c(s0), _c(re), _c(im)
    return lib().orc_s2_energy(len(v), _p(v), _p(s0), _p(re), _p(im), int(N), shift)


def s2_forces(s, a, k, v, s0, re, im, nthreads=0):
    """ShiftedCCPotential_S2::Force for every particle."""
    s, a, k, v, s0, re, im = _c(s), _c(a), _c(k), _c(v), _c(s0), _c(re), _c(im)
    N, d = s.shape
    M = k.shape[0]
    F = np.zeros((N, d))
    if N == 0 or M == 0:
        return F
    lib().orc_s2_forces(d, N, _p(s), _p(a), M, _p(k), _p(v), _p(s0), _p(re), _p(im), _p(F),
                        nthreads or lib().orc_max_threads())
    return F


def softcore_energy(s, a, sigma, e0, nthreads=0):
    s, a = _c(s), _c(a)
    N, d = s.shape
    return lib().orc_softcore_energy(d, N, _p(s), _p(a), sigma, e0, nthreads or lib().orc_max_threads())


def softcore_forces(s, a, sigma, e0, nthreads=0):
    s, a = _c(s), _c(a)
    N, d = s.shape
    F = np.zeros((N, d))
    lib().orc_softcore_forces(d, N, _p(s), _p(a), sigma, e0, _p(F), nthreads or lib().orc_max_threads())
    return F


def total_energy_forces(s, a, k, v, shift=0.0, sigma=0.0, e0=0.0, nthreads=0):
    s, a, k, v = _c(s), _c(a), _c(k), _c(v)
    N, d = s.shape
    M = k.shape[0]
    F = np.zeros((N, d))
    E = lib().orc_total_energy_forces(d, N, _p(s), _p(a), M, _p(k if M else np.zeros((1, d))),
                                      _p(v if M else np.zeros(1)), shift, sigma, e0, _p(F),
                                      nthreads or lib().orc_max_threads())
    return E, F


def contact_radii(s, a, radii, rc, e0, nthreads=0):
    """MultiStealthyPotential's contact term: per-particle energy slots (pairs i < j, before e0) and e0 x forces."""
    s, a, radii = _c(s), _c(a), _c(radii)
    N, d = s.shape
    slots = np.zeros(max(N, 1))
    F = np.zeros((max(N, 1), d))
    lib().orc_contact_radii(d, N, _p(s), _p(a), _p(radii), rc, e0, _p(slots), _p(F),
                            nthreads or lib().orc_max_threads())
    return slots[:N], F[:N]


def multi_energy_forces(s, a, species_idx, species_k, species_v, radii, rc, u0=1.0, e0=1.0, repulsion=True,
                        nthreads=0):
    """MultiStealthyPotential Energy() and AllForce(): species_idx[s] lists the species' particles in its own
    order, species_k[s] / species_v[s] its constraints.  Returns (E, F, species energies)."""
    s, a, radii = _c(s), _c(a), _c(radii)
    N, d = s.shape
    S = len(species_idx)
    idx_off = np.zeros(S + 1, dtype=np.int64)
    k_off = np.zeros(S + 1, dtype=np.int64)
    for q in range(S):
        idx_off[q + 1] = idx_off[q] + len(species_idx[q])
        k_off[q + 1] = k_off[q] + len(species_v[q])
    idx = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.int64) for x in species_idx] +
                                              [np.zeros(1, dtype=np.int64)]))
    k = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, d) for x in species_k] +
                                            [np.zeros((1, d))]))
    v = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in species_v] +
                                            [np.zeros(1)]))
    F = np.zeros((max(N, 1), d))
    ES = np.zeros(max(S, 1))
    ip = C.POINTER(C.c_int64)
    E = lib().orc_multi_energy_forces(d, N, _p(s), _p(a), S, idx_off.ctypes.data_as(ip), idx.ctypes.data_as(ip),
                                      k_off.ctypes.data_as(ip), _p(k), _p(v), _p(radii if N else np.zeros(1)),
                                      rc, u0, e0, int(bool(repulsion)), _p(F), _p(ES),
                                      nthreads or lib().orc_max_threads())
    return E, F[:N], ES[:S]


def relax_emin(s, a, k, v, shift=0.0, sigma=0.0, e0=0.0, maxstep=10000, Emin=-1e10, mf=40, nthreads=0):
    """RelaxStructure_NLOPT_Emin (Emin > -1e10) or RelaxStructure_NLOPT (Emin = -1e10) driven by the
    reference's LuksanPLIS (strict NLopt order).  Returns (E_final, s_final, evaluations, plis_status)."""
    s = _c(s).copy()
    a, k, v = _c(a), _c(k), _c(v)
    N, d = s.shape
    M = k.shape[0]
    plis = C.cast(ref().ref_luksan_plis, C.c_void_p)
    ev = C.c_int(0)
    st = C.c_int(0)
    E = lib().orc_relax_emin(d, N, _p(s), _p(a), M, _p(k), _p(v), shift, sigma, e0, maxstep, Emin, mf, plis,
                             nthreads or lib().orc_max_threads(), C.byref(ev), C.byref(st))
    return E, s, ev.value, st.value


def ref_plis(fun, x0, stopval=-1e10, maxeval=0, mf=40, fast=0):
    """Run the reference LuksanPLIS on a Python objective fun(x) -> (f, g).  Returns (x, minf, status, evals, trace)
    where trace lists every evaluated (x, f)."""
    x = _c(x0).copy()
    n = len(x)
    trace = []

    @OBJECTIVE
    def obj(nn, xp, gp, data):
        xv = np.ctypeslib.as_array(xp, shape=(nn,)).copy()
        f, g = fun(xv)
        if gp:
            ga = np.ctypeslib.as_array(gp, shape=(nn,))
            ga[:] = g
        trace.append((xv, f))
        return f

    minf = C.c_double(0)
    ev = C.c_int(0)
    st = ref().ref_luksan_plis(n, obj, None, _p(x), C.byref(minf), stopval, maxeval, mf, C.byref(ev), fast)
    return x, minf.value, st, ev.value, trace
