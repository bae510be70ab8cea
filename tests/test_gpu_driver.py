"""GetCCO (apps/CCOptimization/main.cpp:135-620) driven by the reference's own input-stream format: ground-state
and MD runs from a text input write the reference's ConfigurationPacks, and every accepted ground state is checked
against the oracle's energy of the same constraints."""
import io

import numpy as np
import pytest

from modules_amd.driver import GetCCO, HyperSphere_Volume, alpha
from modules_amd.formats import ConfigurationPack
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def test_ground_mode_from_an_input_stream(tmp_path):
    save = str(tmp_path / "run")
    # dim K1a K2a S0 val sigma phi threads N numConfig init savename mode, then the ground commands
    text = f"2 0.0 0.8 0 1.0 0.3 0.4 4 60 3 random {save} ground tolerance 1e-14 maxsteps 2000 run"
    out = io.StringIO()
    res = GetCCO(io.StringIO(text), seed=3, out=out)
    assert "chi = " in out.getvalue()
    succ = ConfigurationPack(save + "_Success")
    assert succ.NumConfig() == 3 and ConfigurationPack(save).NumConfig() >= 3
    # the constraint set of the run: GetKs(K2/a) minus |k| <= K1/a, flat weights
    a = (0.4 / HyperSphere_Volume(2, 1.0)) ** 0.5
    for i in range(3):
        c = succ.GetConfig(i)
        k, _ = po.get_ks(c.basis, 0.8 / a)
        E, _ = po.total_energy_forces(c.relative, c.basis, k, np.ones(len(k)), 0.0, 0.3, 1.0)
        assert E < 1e-12
    assert len(res["success"]) == 3


def test_md_mode_from_an_input_stream(tmp_path):
    save = str(tmp_path / "md")
    text = f"2 0.0 0.8 0 1.0 0.3 0.4 4 40 2 random {save} MD timestep 0.01 samplesteps 20 temperature 1e-4 " \
           f"numequilibration 2 run"
    res = GetCCO(io.StringIO(text), seed=1, out=io.StringIO())
    assert len(res) == 2 and ConfigurationPack(save).NumConfig() == 2


def test_overlap_function_closed_forms():
    assert alpha(3, 0.0) == 1.0 and alpha(3, 1.0) == 0.0 and alpha(2, 1.2) == 0.0
    assert alpha(2, 0.5) == pytest.approx(2 / np.pi * (np.arccos(0.5) - 0.5 * np.sqrt(0.75)))
    assert HyperSphere_Volume(3, 2.0) == pytest.approx(4 / 3 * np.pi * 8)


def _multi_text(save, mode_cmds, region="chi", species=((30, 0.0, 0.15), (30, 0.0, 0.15))):
    # random packing: dim, two radius classes (30 spheres of 1.0, 30 of 1.4), phi = 0.3; contact strength 1;
    # per species: count, S0, region word, lower, upper; then threads, save name, configurations, mode
    head = "random 2 2 30 1.0 30 1.4 0.3 1.0 " + str(len(species)) + " "
    body = " ".join(f"{n} {s0} {region} 0 {up}" for n, s0, up in species)
    return f"{head}{body} 4 {save} {mode_cmds}"


def test_multi_ground_mode_from_an_input_stream(tmp_path):
    from modules_amd.driver import GetMultiCCO
    from modules_amd.potential import getK
    save = str(tmp_path / "mix")
    text = _multi_text(save, "3 ground tolerance 1e-14 maxsteps 4000 export txt run")
    out = io.StringIO()
    status, res = GetMultiCCO(io.StringIO(text), seed=2, out=out)
    print(out.getvalue()[-400:], flush=True)
    assert status == 0
    succ = ConfigurationPack(save + "_Success")
    n_ok = succ.NumConfig()
    assert n_ok == len(res["success"]) >= 2
    # radii scaled to phi = 0.3 (main.cpp:1125-1140), constraints 0 < |k| <= getK(0.15, 1, 2) for both species
    factor = (0.3 / (np.pi * (30 * 1.0 + 30 * 1.96) / 60)) ** 0.5
    radii = np.array([factor] * 30 + [1.4 * factor] * 30)
    idx = [np.arange(30), np.arange(30, 60)]
    K = getK(0.15, 1.0, 2)
    for i in range(n_ok):
        c = succ.GetConfig(i)
        k, _ = po.get_ks(c.basis, K)
        E, _, _ = po.multi_energy_forces(c.relative, c.basis, idx, [k, k], [np.ones(len(k))] * 2, radii,
                                         2.01 * radii.max(), 1.0, 1.0)
        assert E < 1e-12
        # the text export of each accepted configuration and the per-species split packs (main.cpp:1591-1611)
        assert (tmp_path / f"mix_Success_{i}.txt").exists()
        for s in range(2):
            cs = ConfigurationPack(f"{save}_s{s}_Success").GetConfig(i)
            np.testing.assert_array_equal(cs.relative, c.relative[idx[s]])


def test_multi_md_mode_follows_the_reference_start_test(tmp_path):
    from modules_amd.driver import GetMultiCCO
    save = str(tmp_path / "mixmd")
    # the start is chosen by the last region word ("chi"): nothing runs
    status, res = GetMultiCCO(io.StringIO(_multi_text(save, "2 MD temperature 1e-4 run")), out=io.StringIO())
    assert status == 0 and res is None and ConfigurationPack(save).NumConfig() == 0
    # an "input" word there stops with the reference's error
    status, _ = GetMultiCCO(io.StringIO(_multi_text(save, "2 MD run", region="input")), out=io.StringIO())
    assert status == 1


def test_multi_input_packing_stops_with_the_reference_error(tmp_path):
    from modules_amd.driver import GetMultiCCO
    from modules_amd.formats import WriteSpherePacking
    from modules_amd.multi import SpherePacking
    rng = np.random.default_rng(4)
    WriteSpherePacking(SpherePacking(np.eye(2) * 6.0, rng.random((36, 2)), np.full(36, 0.3)), tmp_path / "start0")
    text = f"input {tmp_path}/start%d 1.0 1 36 0 chi 0 0.2 4 {tmp_path}/r 2 ground run"
    out = io.StringIO()
    status, _ = GetMultiCCO(io.StringIO(text), out=out)
    assert status == 1 and "N_load" in out.getvalue()
