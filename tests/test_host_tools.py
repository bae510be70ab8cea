"""Host-side pieces that need no GPU: the bench's one-GPU contract, the serving request stream, the work-aware
parts of the Python mirror."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout.strip() == ""


def test_bench_help_lists_the_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0
    for w in ("batch256", "seeds64", "single", "large2d", "mixed2d", "serve"):
        assert w in r.stdout


def test_batch_draw_is_seeded_and_within_its_ranges():
    from modules_amd.serving import make_batch
    a = make_batch()
    b = make_batch()
    assert [(r.N, r.d, r.chi, r.seed) for r in a] == [(r.N, r.d, r.chi, r.seed) for r in b]
    assert len(a) == 256
    assert all(100 <= r.N <= 4000 and r.d in (2, 3) and 0.2 <= r.chi <= 0.45 for r in a)
    assert 90 < sum(r.d == 2 for r in a) < 166
    # log-uniform N: about half of the batch below the geometric mean sqrt(100 * 4000) ~ 632
    assert 90 < sum(r.N < 632 for r in a) < 166
    # the first draw follows numpy's legacy MT19937 stream of the master seed
    rng = np.random.RandomState(20261203)
    assert a[0].N == int(round(np.exp(rng.uniform(np.log(100), np.log(4000)))))


def test_request_stream_is_seeded_and_within_its_ranges():
    from modules_amd.serving import make_stream
    a = make_stream(200, rate=20.0, seed=5)
    b = make_stream(200, rate=20.0, seed=5)
    assert [(r.arrival, r.N, r.d, r.chi, r.seed) for r in a] == [(r.arrival, r.N, r.d, r.chi, r.seed) for r in b]
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    assert 5.0 < arr[-1] < 15.0  # 200 arrivals at 20/s
    assert all(100 <= r.N <= 50000 and r.d in (2, 3) and 0.1 <= r.chi <= 0.49 for r in a)
    assert all(r.sigma == 0.4 and r.epsilon == 1.0 for r in a)
    med = np.median([r.N for r in a])
    assert 700 < med < 1400  # lognormal around a median of 1000


def test_potential_terms_follow_the_constraints():
    import ctypes as C
    from modules_amd import _lib
    h = C.c_void_p()
    if _lib.lib().ccgpu_potential_create(2, C.byref(h)) != 0:
        import pytest
        pytest.skip("potential objects need a GPU here")
    _lib.lib().ccgpu_potential_destroy(h)
    from modules_amd import RepulsiveCCPotential
    from modules_amd.optimize import _potential_terms
    pot = RepulsiveCCPotential(2, 2.0, 0.4)
    pot.AddConstraint([1.0, 0.0], [0.5])
    pot.AddConstraint([0.0, 2.0], [1.5])
    pot.Shift = 0.25
    k, V, shift, e0, sigma = _potential_terms(pot)
    assert np.array_equal(k, [[1.0, 0.0], [0.0, 2.0]]) and np.array_equal(V, [0.5, 1.5])
    assert (shift, e0, sigma) == (0.25, 2.0, 0.4)


@pytest.mark.parametrize("seed", range(6))
def test_varbox_kpoints_match_the_oracle(seed):
    """The library's variable-cell k-point list (host entry) equals the oracle's box enumeration: same points, same
    coefficients, same order, on random skewed cells in 2D and 3D (some strongly sheared)."""
    from modules_amd.varbox import KPoints
    from oracle import varbox as ovb
    rng = np.random.default_rng(seed)
    d = 2 + seed % 2
    a = np.diag(rng.uniform(5.0, 12.0, d))
    a[np.tril_indices(d, -1)] = rng.uniform(-3.0, 3.0, d * (d - 1) // 2)
    if seed >= 4:
        a[1, 0] = 25.0  # strong shear: the bounding box is much larger than the ellipsoid
    rK = rng.uniform(1.0, 2.5)
    k, n = KPoints(a, rK)
    ko, no = ovb.kpoints(a, rK)
    assert len(k) == len(ko) > 0
    assert np.array_equal(n, no)
    assert np.allclose(k, ko, rtol=0, atol=1e-14)


def test_cost_gate_fit_recovers_known_constants():
    # the refit of the gate's direct part: timings made from known constants come back, and the decision count
    # compares the fitted direct time (with the margin) against the NUFFT's prediction and the timed winner
    from modules_amd.fit_cost_gate import MARGIN, fit
    truth = {1: (48.0, 4.0e-7), 2: (42.0, 5.0e-7), 3: (44.0, 6.0e-7)}
    rows = []
    for d, (a, b) in truth.items():
        for N, M in [(300, 100), (2000, 900), (20000, 9000), (100000, 40000), (400000, 160000)]:
            t = a + b * N * M
            # a NUFFT that the gate should take exactly when the direct time clears the margin
            pn = t / (1.5 if N >= 100000 else 0.5)
            rows.append({"N": N, "d": d, "chi": 0.3, "M": M, "direct_us": t, "nufft_us": pn, "pred_nufft_us": pn})
    res = fit(rows)
    for d, (a, b) in truth.items():
        assert abs(res[d]["call_us"] - a) <= 1e-6 * a and abs(res[d]["pair_us"] - b) <= 1e-9 * b
        assert res[d]["max_rel_err"] < 1e-9
        assert res[d]["decisions_agree"] == res[d]["decisions"] == 5
    # the script's margin is the library's
    import re
    hdr = open(os.path.join(ROOT, "modules_amd", "csrc", "cc_runtime.h")).read()
    assert float(re.search(r"kNufftMargin = ([0-9.]+);", hdr).group(1)) == MARGIN


def test_sincos_table_header_is_the_generator_output_and_correctly_rounded():
    """The kernels' 256-entry table: the committed header is what its generator writes, and every entry is the
    double nearest to cos / sin(2 pi t / 256), checked independently in 64-bit-mantissa arithmetic (where the exact
    value is 0, +-1 or +-sqrt(1/2) against the closed form instead)."""
    import importlib.util
    import re

    path = os.path.join(ROOT, "modules_amd", "csrc", "gen_sincos_tab256.py")
    spec = importlib.util.spec_from_file_location("gen_sincos_tab256", path)
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    committed = open(os.path.join(ROOT, "modules_amd", "csrc", "cc_tab256.h")).read()
    assert committed == gen.header()
    vals = [float.fromhex(x) for x in re.findall(r"(-?0x[0-9a-f.]+p[+-]\d+)", committed)]
    tab = np.array(vals).reshape(256, 2)
    t = np.arange(256)
    if np.finfo(np.longdouble).nmant >= 63:
        pi = np.longdouble("3.14159265358979323846264338327950288")
        ang = np.longdouble(2) * pi * t.astype(np.longdouble) / np.longdouble(256)
        c = np.cos(ang).astype(np.float64)
        s = np.sin(ang).astype(np.float64)
        exact = t % 64 == 0  # 0 and +-1, where the long-double values carry a residue of the rounded pi
        assert np.array_equal(tab[~exact, 0], c[~exact]) and np.array_equal(tab[~exact, 1], s[~exact])
    q = t // 64
    assert np.array_equal(tab[t % 64 == 0], np.array([[1.0, 0.0], [0.0, 1.0], [-1.0, 0.0], [0.0, -1.0]])[q[t % 64 == 0]])
    r2 = float.fromhex("0x1.6a09e667f3bcdp-1")  # sqrt(1/2) rounded to nearest
    assert np.array_equal(np.abs(tab[t % 64 == 32]), np.full((4, 2), r2))
    # a planted one-ulp error in one entry is caught by the independent check
    bad = tab.copy()
    bad[37, 0] = np.nextafter(bad[37, 0], 2.0)
    if np.finfo(np.longdouble).nmant >= 63:
        assert not np.array_equal(bad[~exact, 0], c[~exact])
