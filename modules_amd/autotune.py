"""Project autotuner for the E+F kernels' tile shapes (segment length, rho kernel variant, rho particle chunk,
force constraint chunk).  Tiles are recorded per (dimension, N, M) shape class in the engine's process-wide
table; potentials and solvers prepared afterwards use them.  A system's tiles depend only on its own shape,
so its results stay the same whatever else shares a batch."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, dp, lib
from .potential import Configuration, GetKs, getK


def autotune(N, d, chi=None, k=None, basis=None, batch=64, repeats=5):
    """Tune tiles for N particles in d dimensions (cube at density 1 unless `basis` is given) with the
    stealthy constraint set of `chi` (or the explicit wavevectors `k`).  Returns the chosen tiles and the
    milliseconds of one batched E+F evaluation of `batch` configurations."""
    if basis is None:
        basis = np.eye(d) * pow(N, 1.0 / d)
    basis = np.ascontiguousarray(basis, dtype=np.float64)
    if k is None:
        k = GetKs(Configuration(basis), getK(chi, 1.0, d))
    k = np.ascontiguousarray(k, dtype=np.float64)
    v = np.ones(len(k))
    seg, rk, rc, fc, ppl = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    ms = C.c_double(0)
    check(lib().ccgpu_autotune(d, N, basis.ctypes.data_as(dp), len(k), k.ctypes.data_as(dp), v.ctypes.data_as(dp),
                               batch, repeats, C.byref(seg), C.byref(rk), C.byref(rc), C.byref(fc), C.byref(ppl),
                               C.byref(ms)))
    return {"segment_len": seg.value, "rho_kernel": rk.value, "rho_chunk": rc.value, "force_chunk": fc.value,
            "force_ppl": ppl.value, "ms_per_batched_ef": ms.value}


def autotune_backend(N, d, chi=None, k=None, basis=None, eps=1e-12, repeats=5):
    """Time one E+F of this shape on the direct sums and through the NUFFT and record the faster for the shape
    class (systems with the "nufft" backend follow the record).  Returns (use_nufft, ms_direct, ms_nufft); ms_nufft
    is None when the constraint set is not eligible."""
    if basis is None:
        basis = np.eye(d) * pow(N, 1.0 / d)
    basis = np.ascontiguousarray(basis, dtype=np.float64)
    if k is None:
        k = GetKs(Configuration(basis), getK(chi, 1.0, d))
    k = np.ascontiguousarray(k, dtype=np.float64)
    v = np.ones(len(k))
    use, md, mn = C.c_int(0), C.c_double(0), C.c_double(0)
    check(lib().ccgpu_autotune_backend(d, N, basis.ctypes.data_as(dp), len(k), k.ctypes.data_as(dp),
                                       v.ctypes.data_as(dp), float(eps), repeats, C.byref(use), C.byref(md),
                                       C.byref(mn)))
    return bool(use.value), md.value, (mn.value if mn.value >= 0 else None)


def clear():
    """Forget every recorded tile shape and backend choice."""
    check(lib().ccgpu_autotune_clear())


def record(N, d, M, segment_len, rho_kernel, rho_chunk, force_chunk=0, force_ppl=2):
    """Record a tile shape for the (d, N, M) shape class without timing."""
    check(lib().ccgpu_autotune_record(d, N, M, segment_len, rho_kernel, rho_chunk, force_chunk, force_ppl))
