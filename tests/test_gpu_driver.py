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
