"""On-disk formats (host only): ConfigurationPack / SpherePackingPack records, the MSPB v1 snapshot, the MSPB archive
and the text export, against the layouts the reference documents and the expectations of its own export test
(apps/CCOptimization/tests/test_export.cpp: the 2D/3D square and skewed cells with three particles, the byte count
36 + 8 D^2 + N (20 + 8 D), the signature, version 1, 4.0 encoded as 00 00 00 00 00 00 10 40, labels after a
reassignment, the header lines and the table).  The expected bytes here are assembled independently with struct
from the documented field lists (core/PeriodicCellList.h:1254-1267, EXPORT_FORMAT.md, MSPB_PACK.md)."""
import struct

import numpy as np
import pytest

from modules_amd import Configuration
from modules_amd.formats import ConfigurationPack, ExportText, MSPB_Pack, MSPBSnapshot, SpherePackingPack
from modules_amd.multi import SpherePacking


def export_case(dim, skewed):
    """The reference test's case: basis 4 + d on the diagonal (1/7 skew), s_id = (i + d + 1)/11, species after
    SetSpeciesOfParticle(0,0), (0,1), (1,1), (2,0), radii 0.125, 0.25, 0.375."""
    a = np.zeros((dim, dim))
    for d in range(dim):
        a[d, d] = 4.0 + d
    if skewed:
        a[1, 0] = 1.0 / 7.0
    s = np.array([[(i + d + 1) / 11.0 for d in range(dim)] for i in range(3)])
    return a, s, np.array([1, 1, 0]), np.array([0.125, 0.25, 0.375])


def expected_mspb(a, s, lab, radii, nspecies, ids=None):
    d = a.shape[0]
    b = b"MSPB\r\n\x1a\n" + struct.pack("<IIQQI", 1, d, len(s), nspecies, 3)
    b += struct.pack("<%dd" % (d * d), *a.reshape(-1))
    for i in range(len(s)):
        b += struct.pack("<QI", i if ids is None else ids[i], lab[i]) + struct.pack("<%dd" % d, *s[i]) + \
            struct.pack("<d", radii[i])
    return b


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("skewed", [False, True])
def test_mspb_matches_the_reference_export_test(dim, skewed):
    a, s, lab, radii = export_case(dim, skewed)
    snap = MSPBSnapshot(a, s, lab, radii, 2)
    data = snap.Write()
    assert len(data) == 36 + 8 * dim * dim + 3 * (20 + 8 * dim)
    assert data[:8] == b"MSPB\r\n\x1a\n" and data[8:12] == b"\x01\0\0\0"
    assert data[36:44] == b"\0\0\0\0\0\0\x10\x40"
    assert data == expected_mspb(a, s, lab, radii, 2)
    assert snap.Write() == data  # repeatable
    back = MSPBSnapshot.Read(data)
    assert np.array_equal(back.basis, a) and np.array_equal(back.relative, s)
    assert list(back.species) == [1, 1, 0] and np.array_equal(back.radii, radii) and back.nspecies == 2
    assert list(back.ids) == [0, 1, 2]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("skewed", [False, True])
def test_text_export_matches_the_reference_export_test(dim, skewed):
    a, s, lab, radii = export_case(dim, skewed)
    text = ExportText(a, s, lab, radii)
    lines = text.split("\n")
    assert lines[0] == f"# dimension {dim}" and lines[1] == "# particles 3"
    for d in range(dim):
        parts = lines[2 + d].split(" ")
        assert parts[:2] == ["#", f"basis_vector-{d}"]
        assert [float(x) for x in parts[2:]] == list(a[d])
    assert lines[2 + dim] == "# coordinates relative" and lines[3 + dim] == "# radius_units cartesian"
    assert lines[4 + dim] == ("# id\tspecies\tsx\tsy\tradius" if dim == 2 else "# id\tspecies\tsx\tsy\tsz\tradius")
    for i in range(3):
        f = lines[5 + dim + i].split("\t")
        assert int(f[0]) == i and int(f[1]) == (1 if i < 2 else 0)
        assert [float(x) for x in f[2:2 + dim]] == list(s[i])  # 17 significant digits: exact round trip
        assert float(f[2 + dim]) == (i + 1) * 0.125
    assert lines[5 + dim + 3] == "" and len(lines) == 5 + dim + 4
    # C++ ostream at max_digits10 prints like %.17g
    assert ("%.17g" % (1.0 / 7.0)) in text if skewed else True


def test_mspb_reader_rejects_what_the_reference_rejects():
    a, s, lab, radii = export_case(2, True)
    good = MSPBSnapshot(a, s, lab, radii, 2).Write()
    bad = {
        "truncated": good[:-1],
        "trailing": good + b"\0",
        "signature": b"MSPC" + good[4:],
        "version": good[:8] + struct.pack("<I", 2) + good[12:],
        "flags": good[:32] + struct.pack("<I", 1) + good[36:],
        "dimension": good[:12] + struct.pack("<I", 5) + good[16:],
        "nan": expected_mspb(a, np.where(np.arange(6).reshape(3, 2) == 3, np.nan, s), lab, radii, 2),
        "duplicate id": expected_mspb(a, s, lab, radii, 2, ids=[0, 1, 1]),
        "label": expected_mspb(a, s, [1, 2, 0], radii, 2),
        "radius": expected_mspb(a, s, lab, [0.125, 0.0, 0.375], 2),
        "singular": expected_mspb(np.array([[1.0, 2.0], [2.0, 4.0]]), s, lab, radii, 2),
    }
    for name, data in bad.items():
        with pytest.raises(ValueError):
            MSPBSnapshot.Read(data)
        assert name
    # an empty snapshot with no species is valid
    empty = MSPBSnapshot(a, np.zeros((0, 2)), [], [], 0)
    assert MSPBSnapshot.Read(empty.Write()).relative.shape == (0, 2)


def test_mspb_archive_layout_and_round_trip(tmp_path):
    prefix = str(tmp_path / "ensemble")
    pack = MSPB_Pack(prefix)
    assert pack.NumConfig() == 0
    snaps = [MSPBSnapshot(*export_case(d, sk), 2 + d) for d, sk in ((2, False), (3, True))]
    for sn in snaps:
        pack.AddConfig(sn)
    raw = open(prefix + ".MSPBPack", "rb").read()
    want = b"MSPP\r\n\x1a\n" + struct.pack("<II", 1, 0)
    for sn in snaps:
        w = sn.Write()
        want += struct.pack("<Q", len(w)) + w
    assert raw == want
    assert MSPB_Pack(prefix).NumConfig() == 2
    assert pack.GetConfig(1).Write() == snaps[1].Write()
    out = str(tmp_path / "copy.mspb")
    pack.ExportFile(0, out)
    assert open(out, "rb").read() == snaps[0].Write()
    with pytest.raises(FileExistsError):
        pack.ExportFile(0, out)
    pack.ImportFile(out)
    assert pack.NumConfig() == 3
    with pytest.raises(IndexError):
        pack.GetConfig(3)
    open(prefix + ".MSPBPack", "ab").write(b"\x05")  # truncated framing
    with pytest.raises(ValueError):
        MSPB_Pack(prefix)


def expected_record(c, names=None, radii=None, cell=None):
    d = c.GetDimension()
    b = struct.pack("<Hd", d, cell) + struct.pack("<%dd" % (d * d), *c.basis.reshape(-1))
    b += struct.pack("<i", c.NumParticle())
    for i in range(c.NumParticle()):
        b += struct.pack("<%dd" % d, *c.relative[i])
        b += struct.pack("<d", radii[i]) if radii is not None else (names[i] if names else b"a\0\0\0")
    return b


def test_configuration_pack_records_index_and_round_trip(tmp_path):
    prefix = str(tmp_path / "run")
    pack = ConfigurationPack(prefix)
    assert pack.NumConfig() == 0
    c1 = Configuration(np.diag([3.0, 4.0]), np.array([[0.1, 0.2], [0.7, 0.9], [0.25, 0.5]]))
    c1.cell_size = 0.4
    c2 = Configuration(np.array([[2.0, 0, 0], [0.3, 2.5, 0], [0.1, 0.2, 3.0]]), np.random.default_rng(1).random((5, 3)))
    c2.names = ["b", "cc", "ddd", "eeee", "a"]
    pack.AddConfig(c1)
    pack.AddConfig(c2)
    r1 = expected_record(c1, cell=0.4)
    r2 = expected_record(c2, names=[b"b\0\0\0", b"cc\0\0", b"ddd\0", b"eee\0", b"a\0\0\0"],
                         cell=0.1 * np.linalg.norm(c2.basis[0]))
    assert open(prefix + ".ConfigPack", "rb").read() == r1 + r2
    assert open(prefix + ".ConfigPackIndex", "rb").read() == struct.pack("<qqq", 2, 0, len(r1))
    again = ConfigurationPack(prefix)
    assert again.NumConfig() == 2
    g = again.GetConfig(1)
    assert np.array_equal(g.basis, c2.basis) and np.array_equal(g.relative, c2.relative)
    assert g.names == [b"b", b"cc", b"ddd", b"eee", b"a"] and g.cell_size == 0.1 * np.linalg.norm(c2.basis[0])
    with pytest.raises(IndexError):
        again.GetConfig(2)
    again.Clear()
    assert ConfigurationPack(prefix).NumConfig() == 0


def test_sphere_packing_pack(tmp_path):
    prefix = str(tmp_path / "spheres")
    p = SpherePacking(np.diag([5.0, 6.0]), np.array([[0.1, 0.3], [0.6, 0.2]]), [0.5, 0.75])
    p.cell_size = 1.5
    pack = SpherePackingPack(prefix)
    pack.AddConfig(p)
    assert open(prefix + ".SpherePack", "rb").read() == expected_record(p, radii=[0.5, 0.75], cell=1.5)
    q = SpherePackingPack(prefix).GetConfig(0)
    assert np.array_equal(q.radii, p.radii) and np.array_equal(q.relative, p.relative)


def mddump_bytes(stage, step, a, cs, s, names, dt, t, xi, q, mass, v, acc):
    """The MD restart file assembled with struct from the reference's writes: main.cpp:1938-1941 (int stage,
    long long step), MD_System.cpp:50-62 (Position record, TimeStep, Time, size_t NumParticle, unsigned short
    Dimension, xi, MassQ, Mass, then Velocity and Acceleration as GeometryVector = double x[4])."""
    d, n = a.shape[0], len(s)
    b = struct.pack("<iq", stage, step) + struct.pack("<Hd", d, cs) + struct.pack("<%dd" % (d * d), *a.reshape(-1))
    b += struct.pack("<i", n)
    for i in range(n):
        b += struct.pack("<%dd" % d, *s[i]) + names[i].ljust(4, b"\0")
    b += struct.pack("<ddQHdd", dt, t, n, d, xi, q) + struct.pack("<%dd" % n, *mass)
    for vec in (v, acc):
        for i in range(n):
            b += struct.pack("<4d", *(list(vec[i]) + [0.0] * (4 - d)))
    return b


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_mddump_layout_and_round_trip(tmp_path, dim):
    # written by the library (ccgpu_mddump_write), compared byte for byte with the documented field list, then
    # read back (ccgpu_mddump_read) exactly
    import ctypes as C
    from modules_amd._lib import check, lib
    from modules_amd.potential import _p
    rng = np.random.default_rng(dim)
    a, _, _, _ = export_case(max(dim, 2), False)
    a = a[:dim, :dim].copy()
    n = 5
    s, v, acc = rng.random((n, dim)), rng.normal(size=(n, dim)), rng.normal(size=(n, dim))
    mass = rng.uniform(0.5, 2.0, n)
    names = [b"a", b"b", b"Cu", b"a", b"xyz"]
    chars = np.frombuffer(b"".join(x.ljust(4, b"\0") for x in names), dtype=np.uint8).copy()
    path = str(tmp_path / "run.MDDump")
    check(lib().ccgpu_mddump_write(path.encode(), 2, 17, dim, 0.4, _p(a), n, _p(s), chars.ctypes.data_as(C.c_void_p),
                                   0.01, 1.25, -0.3, 7.5, _p(mass), _p(v), _p(acc)))
    raw = open(path, "rb").read()
    assert raw == mddump_bytes(2, 17, a, 0.4, s, names, 0.01, 1.25, -0.3, 7.5, mass, v, acc)
    st, sp, d, N = C.c_int32(0), C.c_int64(0), C.c_int(0), C.c_int64(0)
    cs, dt, t, xi, q = (C.c_double(0) for _ in range(5))
    basis = np.zeros(16)
    s2, v2, a2, m2 = np.zeros((n, dim)), np.zeros((n, dim)), np.zeros((n, dim)), np.zeros(n)
    c2 = np.zeros(4 * n, dtype=np.uint8)
    check(lib().ccgpu_mddump_read(path.encode(), C.byref(st), C.byref(sp), C.byref(d), C.byref(cs), _p(basis),
                                  C.byref(N), n, _p(s2), c2.ctypes.data_as(C.c_void_p), C.byref(dt), C.byref(t),
                                  C.byref(xi), C.byref(q), _p(m2), _p(v2), _p(a2)))
    assert (st.value, sp.value, d.value, N.value) == (2, 17, dim, n)
    assert (cs.value, dt.value, t.value, xi.value, q.value) == (0.4, 0.01, 1.25, -0.3, 7.5)
    assert np.array_equal(basis[:dim * dim], a.reshape(-1))
    for x, y in ((s2, s), (v2, v), (a2, acc), (m2, mass)):
        assert np.array_equal(x, y)
    assert bytes(c2) == bytes(chars)


def test_mddump_reader_rejects_damaged_files(tmp_path):
    import ctypes as C
    from modules_amd._lib import CCGPUError, check, lib
    from modules_amd.potential import _p
    a = np.eye(2) * 3.0
    s = np.array([[0.1, 0.2], [0.3, 0.4]])
    good = mddump_bytes(1, 0, a, 0.3, s, [b"a", b"a"], 0.01, 0.0, 0.0, 0.0, [1.0, 1.0], np.zeros((2, 2)),
                        np.zeros((2, 2)))
    bad_count = bytearray(good)
    off = 4 + 8 + 2 + 8 + 32 + 4 + 2 * 20 + 16  # the NumParticle field after the record, TimeStep and Time
    bad_count[off:off + 8] = struct.pack("<Q", 3)
    cases = {"short": good[:-1], "long": good + b"\0", "count": bytes(bad_count), "empty": b""}
    for name, data in cases.items():
        p = tmp_path / f"{name}.MDDump"
        p.write_bytes(data)
        st, sp, d, N = C.c_int32(0), C.c_int64(0), C.c_int(0), C.c_int64(0)
        cs, dt, t, xi, q = (C.c_double(0) for _ in range(5))
        with pytest.raises(CCGPUError):
            check(lib().ccgpu_mddump_read(str(p).encode(), C.byref(st), C.byref(sp), C.byref(d), C.byref(cs),
                                          _p(np.zeros(16)), C.byref(N), 0, None, None, C.byref(dt), C.byref(t),
                                          C.byref(xi), C.byref(q), None, None, None))
    p = tmp_path / "good.MDDump"
    p.write_bytes(good)
    st, sp, d, N = C.c_int32(0), C.c_int64(0), C.c_int(0), C.c_int64(0)
    cs, dt, t, xi, q = (C.c_double(0) for _ in range(5))
    check(lib().ccgpu_mddump_read(str(p).encode(), C.byref(st), C.byref(sp), C.byref(d), C.byref(cs),
                                  _p(np.zeros(16)), C.byref(N), 0, None, None, C.byref(dt), C.byref(t), C.byref(xi),
                                  C.byref(q), None, None, None))
    assert N.value == 2 and st.value == 1
