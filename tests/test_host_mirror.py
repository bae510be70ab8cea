"""Host-side mirror pieces that run without a GPU: RandomGenerator streams and coordinate conversions."""
import numpy as np

from modules_amd import Configuration, RandomConfiguration
from modules_amd.potential import RandomGenerator
from oracle import pyoracle as po


def test_random_generator_is_the_mt19937_stream():
    gen = RandomGenerator(77)
    first = np.array([gen.RandomDouble() for _ in range(5)])
    rest = gen.RandomDoubles(10000)  # crosses the first buffer refill
    u = po.mt_sequence(77, 10005).astype(np.float64) / 4294967296.0
    assert np.array_equal(np.concatenate([first, rest]), u)


def test_random_configuration_continues_a_generator():
    # RandomConfiguration(seed) equals the first N*d draws; a generator continues where the last call stopped
    N, d = 40, 3
    gen = RandomGenerator(5)
    c0 = RandomConfiguration(gen, N, d)
    c1 = RandomConfiguration(gen, N, d)
    assert np.array_equal(c0.relative, RandomConfiguration(5, N, d).relative)
    both = RandomConfiguration(5, 2 * N, d).relative
    assert np.array_equal(c1.relative, both[N:])
    assert np.array_equal(c0.basis, np.eye(d) * N ** (1.0 / d))


def test_cartesian_relative_round_trip():
    basis = np.array([[1.3, 0.0, 0.0], [0.45, 0.9, 0.0], [0.1, -0.2, 1.7]])
    c = Configuration(basis, np.random.default_rng(3).random((5, 3)))
    b = np.array([c.GetReciprocalBasisVector(i) for i in range(3)])
    assert np.allclose(basis @ b.T, 2 * np.pi * np.eye(3), atol=1e-13)
    for i in range(5):
        r = c.GetCartesianCoordinates(i)
        assert np.allclose(c.CartesianCoord2RelativeCoord(r), c.relative[i], atol=1e-14)
