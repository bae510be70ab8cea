"""The allocation-free evaluation (ccgpu_energy_and_all_force_noalloc) against the bound-potential entry point
(ccgpu_potential_energy_and_all_force, CombinedEnergyForce::EnergyAndAllForce, potential/CombinedEnergyForce.h:
17-28) through the C ABI: the same bits on consecutive calls that shrink, on the lattice and the general path, and a
clear error for an evaluation larger than the scratch array."""
import ctypes as C

import numpy as np
import pytest

from modules_amd import GetKs, RandomConfiguration, getK
from modules_amd._lib import ERR_MEMORY, CCGPUError, check, lib
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def bound_potential(d, basis, rel, k, v):
    """Energy and forces through a potential bound to the configuration (the existing entry point)."""
    L = lib()
    h = C.c_void_p()
    check(L.ccgpu_potential_create(d, C.byref(h)))
    try:
        check(L.ccgpu_potential_set_constraints(h, len(v), _p(k), _p(v)))
        check(L.ccgpu_potential_set_configuration(h, _p(basis), len(rel), _p(rel)))
        E = C.c_double()
        F = np.empty_like(rel)
        check(L.ccgpu_potential_energy_and_all_force(h, C.byref(E), _p(F)))
        return E.value, F
    finally:
        check(L.ccgpu_potential_destroy(h))


def noalloc(d, basis, rel, k, v):
    E = C.c_double()
    F = np.full_like(rel, np.nan)
    check(lib().ccgpu_energy_and_all_force_noalloc(d, _p(basis), len(rel), _p(rel), len(v), _p(k), _p(v),
                                                   C.byref(E), _p(F)))
    return E.value, F


def case(N, d, chi, seed, general=False):
    cfg = RandomConfiguration(seed, N, d)
    k = np.ascontiguousarray(GetKs(cfg, getK(chi, 1.0, d)), dtype=np.float64)
    if general:  # off the reciprocal lattice: the Cartesian-wavevector kernels
        k = k * (1.0 + 1e-3 * np.random.RandomState(seed).rand(*k.shape))
    v = np.random.RandomState(seed + 1).uniform(0.5, 1.5, len(k))
    basis = np.ascontiguousarray(cfg.basis, dtype=np.float64)
    rel = np.ascontiguousarray(cfg.relative, dtype=np.float64)
    return d, basis, rel, k, v


def same_bits(c):
    Eb, Fb = bound_potential(*c)
    En, Fn = noalloc(*c)
    assert En == Eb and np.array_equal(Fn, Fb), (En, Eb, np.abs(Fn - Fb).max())
    return En, Fn


def test_noalloc_matches_the_bound_potential_on_shrinking_calls():
    big = case(3000, 3, 0.42, 11)
    E1, F1 = same_bits(big)
    small = case(400, 2, 0.3, 12)  # the second call is smaller: nothing of the first may leak into it
    E2, F2 = same_bits(small)
    # and the shared array keeps no state that changes a repeated call
    assert noalloc(*small)[0] == E2
    # the oracle agrees to the scale-aware 1e-12 of the potential tests
    d, basis, rel, k, v = small
    Eo, Fo = po.total_energy_forces(rel, basis, k, v)
    scale = len(k) * len(rel) ** 2 / abs(np.linalg.det(basis))
    assert abs(Eo - E2) <= 1e-12 * scale and np.abs(Fo - F2).max() <= 1e-12 * scale


def test_noalloc_general_path_and_oversize_rejection():
    same_bits(case(700, 3, 0.35, 21, general=True))
    # 1.2 M particles in 2D need more than the 64 MiB array: refused before any kernel runs
    d, basis, _, k, v = case(10, 2, 0.3, 22)
    rel = np.random.RandomState(5).rand(1_200_000, 2)
    with pytest.raises(CCGPUError) as ei:
        noalloc(d, basis, rel, k, v)
    assert ei.value.code == ERR_MEMORY and "scratch" in str(ei.value)
    # the array is usable again at once
    same_bits(case(200, 2, 0.25, 23))
