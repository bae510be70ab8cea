"""The C-ABI library loads without a GPU and exports every entry point include/ccgpu.h declares, and the
host-only entry points (lattice helpers, seeded configurations) agree with the oracle bit for bit."""
import os
import re

import numpy as np

from modules_amd import _lib
from modules_amd import GetKs, RandomConfiguration, getK
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "ccgpu.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|double|const char \*)\s*(ccgpu_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported_and_bound():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_no_device_is_reported_cleanly():
    import ctypes as C
    h = C.c_void_p()
    code = _lib.lib().ccgpu_potential_create(2, C.byref(h))
    if code != 0:  # a CPU-only host: creation fails loudly with the device status
        assert code == _lib.ERR_DEVICE
        assert b"device" in _lib.lib().ccgpu_last_error()
    else:
        _lib.lib().ccgpu_potential_destroy(h)


def test_product_random_configuration_equals_oracle():
    for seed, N, d in ((0, 200, 2), (1, 2000, 3), (4294967295, 17, 3)):
        cfg = RandomConfiguration(seed, N, d)
        assert np.array_equal(cfg.relative, po.make_config(seed, N, d))
        assert cfg.basis[0, 0] == pow(N, 1.0 / d)


def test_product_get_ks_equals_oracle_bitwise():
    cases = [(np.eye(2) * np.sqrt(200), getK(0.3, 1, 2)), (np.eye(3) * 2000 ** (1 / 3), getK(0.45, 1, 3)),
             (np.eye(2) * np.sqrt(20000), getK(0.4, 1, 2)), (np.array([[1.3, 0.0], [0.45, 0.9]]) * 6, 5.0),
             (np.array([[3.0, 0.0, 0.0], [0.5, 2.5, 0.0], [0.2, 0.3, 4.0]]), 6.0), (np.array([[10.0]]), 3.0)]
    for a, K in cases:
        k, n = GetKs(a, K, return_indices=True)
        ko, no = po.get_ks(a, K)
        assert np.array_equal(k, ko) and np.array_equal(n, no)
    assert getK(0.45, 1.0, 3) == po.getK(0.45, 1.0, 3)


def test_product_reciprocal_basis_and_volume_equal_oracle():
    import ctypes as C
    for a in (np.eye(3) * 7.1, np.array([[1.3, 0.0], [0.45, 0.9]]),
              np.array([[3.0, 0.0, 0.0], [0.5, 2.5, 0.0], [0.2, 0.3, 4.0]])):
        d = len(a)
        b = np.zeros((d, d))
        ac = np.ascontiguousarray(a)
        _lib.check(_lib.lib().ccgpu_reciprocal_basis(d, ac.ctypes.data_as(_lib.dp), b.ctypes.data_as(_lib.dp)))
        assert np.array_equal(b, po.reciprocal_basis(a))
        assert _lib.lib().ccgpu_cell_volume(d, ac.ctypes.data_as(_lib.dp)) == po.volume(a)
