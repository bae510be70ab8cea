"""Host parts of the drivers: the token conversions of the multi-species input stream and the SpherePacking text
files of core/ConfigurationIO.cpp:763-908 (no GPU)."""
import numpy as np

from modules_amd.driver import _atof, _atoi
from modules_amd.formats import ReadSpherePacking, WriteSpherePacking
from modules_amd.multi import SpherePacking


def test_atoi_atof_take_the_leading_number():
    assert _atoi("30") == 30 and _atoi("3.7") == 3 and _atoi("-2x") == -2 and _atoi("abc") == 0
    assert _atof("1.4") == 1.4 and _atof("2e-1z") == 0.2 and _atof(".5") == 0.5 and _atof("x") == 0.0


def test_sphere_packing_text_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    basis = np.array([[5.0, 0.0, 0.0], [0.5, 4.0, 0.0], [0.0, 0.25, 6.0]])
    pck = SpherePacking(basis, rng.random((17, 3)), rng.random(17))
    WriteSpherePacking(pck, tmp_path / "p")
    back = ReadSpherePacking(tmp_path / "p")
    np.testing.assert_array_equal(back.basis, basis)
    np.testing.assert_array_equal(back.radii, pck.radii)
    np.testing.assert_allclose(back.relative, pck.relative, rtol=0, atol=1e-14)


def test_sphere_packing_text_layout(tmp_path):
    # the writer's layout: the dimension, basis rows of four columns, then four coordinates and the radius per
    # sphere; coordinates past d are ignored, a sphere outside the cell wraps, an incomplete last row is dropped
    (tmp_path / "q.txt").write_text("2\n4\t0\t0\t0\n0\t2\t0\t0\n"
                                    "1\t1\t7\t7\t0.5\n5\t-1\t0\t0\t0.25\n3\t1\t0\n")
    q = ReadSpherePacking(tmp_path / "q")
    np.testing.assert_array_equal(q.basis, [[4, 0], [0, 2]])
    np.testing.assert_array_equal(q.radii, [0.5, 0.25])
    np.testing.assert_allclose(q.relative, [[0.25, 0.5], [0.25, 0.5]])
    text = (tmp_path / "q.txt").read_text()
    WriteSpherePacking(q, tmp_path / "r")
    assert (tmp_path / "r.txt").read_text().splitlines()[1].split("\t") == ["4.0", "0.0", "0.0", "0.0"]
    assert text.startswith("2\n")
