"""On-disk formats of the reference's CC tools, over the library's host-side readers and writers (include/ccgpu.h).

  * ConfigurationPack / SpherePackingPack   core/PeriodicCellList.h:1598-1640, core/PeriodicCellList.cpp:124-204,
                                            core/SpherePacking.h:215-251 (prefix.ConfigPack[Index],
                                            prefix.SpherePack[Index]; records in PeriodicCellList::WriteBinary layout)
  * MSPBSnapshot                            apps/CCOptimization/EXPORT_FORMAT.md (MSPB v1)
  * MSPB_Pack                               apps/CCOptimization/MSPB_PACK.md (prefix.MSPBPack)
The byte layouts are the reference's, so files written here are read by its tools and the other way round.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import CCGPUError, ERR_RANGE, check, ip32, lib, u32p, u64p
from .potential import Configuration, _p


def _default_cell_size(c):
    # GetUnitCubicBox(d, 0.1) then Rescale(L) (core/PeriodicCellList.cpp:106-118, core/PeriodicCellList.h:1201-1210)
    cs = getattr(c, "cell_size", None)
    return float(cs) if cs else 0.1 * float(np.linalg.norm(c.basis[0]))


class ConfigurationPack:
    """A pack of configurations in prefix.ConfigPack and prefix.ConfigPackIndex (append-only, indexed)."""

    _spheres = 0

    def __init__(self, prefix=None):
        self.prefix = None
        if prefix is not None:
            self.Open(prefix)

    def Open(self, prefix):
        self.prefix = str(prefix)

    def _b(self):
        if self.prefix is None:
            raise RuntimeError("ConfigurationPack: no file prefix (Open first)")
        return os.fsencode(self.prefix)

    def NumConfig(self):
        n = C.c_int64(0)
        check(lib().ccgpu_pack_count(self._b(), self._spheres, C.byref(n)))
        return n.value

    def _chars(self, c):
        names = getattr(c, "names", None)
        if names is None:
            return None
        raw = b"".join((n.encode() if isinstance(n, str) else bytes(n))[:3].ljust(4, b"\0") for n in names)
        return np.ascontiguousarray(np.frombuffer(raw, dtype=np.uint8))

    def AddConfig(self, c):
        d = c.GetDimension()
        rel = np.ascontiguousarray(c.relative, dtype=np.float64)
        chars = self._chars(c)
        check(lib().ccgpu_pack_add(self._b(), self._spheres, d, _default_cell_size(c), _p(np.ascontiguousarray(c.basis)),
                                   rel.shape[0], _p(rel), None if chars is None else chars.ctypes.data_as(C.c_void_p)))

    def _read(self, i):
        d = C.c_int(0)
        cs = C.c_double(0)
        basis = np.zeros(16)
        n = C.c_int64(0)
        status = lib().ccgpu_pack_get(self._b(), self._spheres, int(i), C.byref(d), C.byref(cs), _p(basis),
                                      C.byref(n), 0, None, None)
        if status == ERR_RANGE:
            raise IndexError("ConfigurationPack::GetConfig: index is out of range")
        check(status)
        rel = np.zeros((max(n.value, 1), d.value))
        cb = 8 if self._spheres else 4
        chars = np.zeros(max(n.value, 1) * cb, dtype=np.uint8)
        check(lib().ccgpu_pack_get(self._b(), self._spheres, int(i), C.byref(d), C.byref(cs), _p(basis), C.byref(n),
                                   n.value, _p(rel), chars.ctypes.data_as(C.c_void_p)))
        dd = d.value
        return basis[: dd * dd].reshape(dd, dd), rel[: n.value], chars[: n.value * cb], cs.value

    def GetConfig(self, i):
        basis, rel, chars, cs = self._read(i)
        c = Configuration(basis, rel)
        c.cell_size = cs
        c.names = [bytes(chars[4 * q: 4 * q + 4]).split(b"\0")[0] for q in range(len(rel))]
        return c

    def Clear(self):
        check(lib().ccgpu_pack_clear(self._b(), self._spheres))


class SpherePackingPack(ConfigurationPack):
    """The SpherePacking counterpart: prefix.SpherePack and prefix.SpherePackIndex, one radius per particle."""

    _spheres = 1

    def _chars(self, c):
        return np.ascontiguousarray(np.asarray(c.radii, dtype=np.float64)).view(np.uint8)

    def GetConfig(self, i):
        from .multi import SpherePacking
        basis, rel, chars, cs = self._read(i)
        p = SpherePacking(basis, rel, chars.view(np.float64) if len(rel) else np.zeros(0))
        p.cell_size = cs
        return p


class MSPBSnapshot:
    """One MSPB v1 snapshot: cell, particle ids, species labels, relative coordinates, radii, declared species count."""

    def __init__(self, basis, relative, species, radii, nspecies, ids=None):
        self.basis = np.array(basis, dtype=np.float64)
        d = self.basis.shape[0]
        self.relative = np.array(relative, dtype=np.float64).reshape(-1, d)
        n = self.relative.shape[0]
        self.species = np.array(species, dtype=np.uint32).reshape(n)
        self.radii = np.array(radii, dtype=np.float64).reshape(n)
        self.ids = np.arange(n, dtype=np.uint64) if ids is None else np.array(ids, dtype=np.uint64).reshape(n)
        self.nspecies = int(nspecies)

    def Write(self):
        """The snapshot's MSPB bytes."""
        d = self.basis.shape[0]
        n = self.relative.shape[0]
        size = C.c_int64(0)
        args = (d, _p(np.ascontiguousarray(self.basis)), n, self.ids.ctypes.data_as(u64p),
                self.species.ctypes.data_as(u32p), _p(np.ascontiguousarray(self.relative)),
                _p(np.ascontiguousarray(self.radii)), self.nspecies)
        status = lib().ccgpu_mspb_encode(*args, None, 0, C.byref(size))
        if status != 0:
            raise ValueError(lib().ccgpu_last_error().decode())
        buf = (C.c_uint8 * size.value)()
        check(lib().ccgpu_mspb_encode(*args, buf, size.value, C.byref(size)))
        return bytes(buf)

    @staticmethod
    def Read(data):
        """Parse and validate MSPB bytes (or a binary file object); ValueError for anything the reference's reader
        rejects."""
        if hasattr(data, "read"):
            data = data.read()
        data = bytes(data)
        buf = (C.c_uint8 * max(len(data), 1)).from_buffer_copy(data.ljust(1, b"\0"))
        d = C.c_int(0)
        basis = np.zeros(16)
        n = C.c_int64(0)
        ns = C.c_uint64(0)
        status = lib().ccgpu_mspb_decode(buf, len(data), C.byref(d), _p(basis), C.byref(n), C.byref(ns), 0, None, None,
                                         None, None)
        if status != 0:
            raise ValueError(lib().ccgpu_last_error().decode())
        N, dd = n.value, d.value
        ids = np.zeros(max(N, 1), dtype=np.uint64)
        lab = np.zeros(max(N, 1), dtype=np.uint32)
        rel = np.zeros((max(N, 1), dd))
        radii = np.zeros(max(N, 1))
        check(lib().ccgpu_mspb_decode(buf, len(data), C.byref(d), _p(basis), C.byref(n), C.byref(ns), N,
                                      ids.ctypes.data_as(u64p), lab.ctypes.data_as(u32p), _p(rel), _p(radii)))
        return MSPBSnapshot(basis[: dd * dd].reshape(dd, dd), rel[:N], lab[:N], radii[:N], ns.value, ids[:N])

    def ToSpherePacking(self):
        from .multi import SpherePacking
        return SpherePacking(self.basis, self.relative, self.radii)


class MSPB_Pack:
    """An ordered ensemble of MSPB snapshots in prefix.MSPBPack (append-only; payload bytes kept unchanged)."""

    def __init__(self, prefix=None):
        self.prefix = None
        if prefix is not None:
            self.Open(prefix)

    def Open(self, prefix):
        self.prefix = str(prefix)
        self.NumConfig()  # a malformed archive is rejected at once

    def NumConfig(self):
        n = C.c_int64(0)
        status = lib().ccgpu_mspbpack_count(os.fsencode(self.prefix), C.byref(n))
        if status != 0:
            raise ValueError(lib().ccgpu_last_error().decode())
        return n.value

    def _add_bytes(self, data):
        buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
        status = lib().ccgpu_mspbpack_add(os.fsencode(self.prefix), buf, len(data))
        if status != 0:
            raise ValueError(lib().ccgpu_last_error().decode())

    def AddConfig(self, snapshot):
        self._add_bytes(snapshot.Write())

    def ImportFile(self, path):
        with open(path, "rb") as f:
            self._add_bytes(f.read())

    def _get_bytes(self, i):
        size = C.c_int64(0)
        status = lib().ccgpu_mspbpack_get(os.fsencode(self.prefix), int(i), None, 0, C.byref(size))
        if status == ERR_RANGE:
            raise IndexError("MSPB_Pack::GetConfig: index is out of range")
        if status != 0:
            raise CCGPUError(status, lib().ccgpu_last_error().decode())
        buf = (C.c_uint8 * size.value)()
        check(lib().ccgpu_mspbpack_get(os.fsencode(self.prefix), int(i), buf, size.value, C.byref(size)))
        return bytes(buf)

    def GetConfig(self, i):
        return MSPBSnapshot.Read(self._get_bytes(i))

    def ExportFile(self, i, path):
        if os.path.exists(path):
            raise FileExistsError(path)
        with open(path, "wb") as f:
            f.write(self._get_bytes(i))

    def Clear(self):
        open(self.prefix + ".MSPBPack", "wb").close()


def ExportText(basis, relative, species, radii):
    """The text table of MultiStealthyPotential::ExportConfiguration (MultiStealthyPotential.cpp:82-131)."""
    basis = np.ascontiguousarray(basis, dtype=np.float64)
    d = basis.shape[0]
    rel = np.ascontiguousarray(relative, dtype=np.float64).reshape(-1, d)
    n = rel.shape[0]
    lab = np.ascontiguousarray(species, dtype=np.int32).reshape(n)
    rad = np.ascontiguousarray(radii, dtype=np.float64).reshape(n)
    size = C.c_int64(0)
    args = (d, _p(basis), n, lab.ctypes.data_as(ip32), _p(rel), _p(rad))
    status = lib().ccgpu_export_text(*args, None, 0, C.byref(size))
    if status != 0:
        raise ValueError(lib().ccgpu_last_error().decode())
    buf = C.create_string_buffer(size.value)
    check(lib().ccgpu_export_text(*args, buf, size.value, C.byref(size)))
    return buf.raw[: size.value].decode()


_MAX_DIMENSION = 4  # core/etc.h:12: every coordinate row of the text files holds four columns


def WriteSpherePacking(pck, OutputName):
    """WriteConfiguration(SpherePacking, name) (core/ConfigurationIO.cpp:763-785): name.txt holding the dimension,
    one row per basis vector and one row per sphere (four Cartesian columns, then the radius), 17 digits."""
    d = pck.GetDimension()
    basis = np.asarray(pck.basis, dtype=np.float64)
    cart = np.asarray(pck.relative, dtype=np.float64).reshape(-1, d) @ basis
    pad = np.zeros(_MAX_DIMENSION)

    def row(v):
        w = pad.copy()
        w[: len(v)] = v
        return "\t".join(repr(float(x)) for x in w)

    lines = [str(d)] + [row(basis[i]) for i in range(d)]
    for i in range(cart.shape[0]):
        lines.append(row(cart[i]) + "\t" + repr(float(pck.radii[i])))
    with open(str(OutputName) + ".txt", "w") as f:
        f.write("\n".join(lines) + "\n")


def ReadSpherePacking(InputName):
    """ReadConfiguration(SpherePacking, name) (core/ConfigurationIO.cpp:865-908): reads name.txt; the first d
    numbers of each basis row are the basis vector, then every five numbers (four coordinates, of which the first
    d are used, and the radius) are one sphere, converted to relative coordinates and wrapped into [0, 1).  A
    trailing incomplete row is dropped, as the reference's end-of-file test drops it."""
    from .multi import SpherePacking
    with open(str(InputName) + ".txt") as f:
        lines = f.read().split("\n")
    head = lines[0].split()
    d = int(head[0])
    rest = head[1:]
    basis = np.zeros((d, d))
    li = 1
    for i in range(d):
        # `ifile >> x` skips blank lines, then the rest of the basis row is ignored
        vals = list(rest)
        rest = []
        while len(vals) < d:
            vals += lines[li].split()
            li += 1
        basis[i] = [float(x) for x in vals[:d]]
    toks = " ".join(lines[li:]).split()
    per = _MAX_DIMENSION + 1
    n = len(toks) // per
    cart = np.zeros((n, d))
    radii = np.zeros(n)
    for p in range(n):
        row = toks[p * per:(p + 1) * per]
        cart[p] = [float(x) for x in row[:d]]
        radii[p] = float(row[_MAX_DIMENSION])
    rel = cart @ np.linalg.inv(basis) if n else np.zeros((0, d))
    rel = rel - np.floor(rel)
    rel = rel - np.floor(rel)
    return SpherePacking(basis, rel, radii)
