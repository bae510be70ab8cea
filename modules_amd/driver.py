"""GetCCO, the single-species driver of apps/CCOptimization/main.cpp:135-620, reading the same input stream.

The reference reads its parameters from standard input token by token (main.cpp:182-249) and then the commands of
the chosen mode until `run` (MD: :345-404, ground: :452-477).  This driver reads the same tokens from a text stream,
builds the same RepulsiveCCPotential (constraints K1/a < |k| < K2/a weighted by the flat, overlap or power-law
tildeV, or the target-S(k) form when S0 != 0, :250-318) and runs CollectiveCoordinateMultiRun (ground) or
CollectiveCoordinateMD (MD) on the engine, writing the reference's ConfigurationPacks under the save name.
"""
from __future__ import annotations

import math
import sys

from .potential import GetKs, RandomConfiguration, RandomGenerator, RepulsiveCCPotential

_BALL = {1: 2.0, 2: math.pi, 3: 4.0 * math.pi / 3.0, 4: math.pi ** 2 / 2.0, 5: 8.0 * math.pi ** 2 / 15.0}


def HyperSphere_Volume(d, R):
    """HyperSphere_Volume (core/etc.cpp:71-74)."""
    return _BALL[d] * R ** d


def alpha(d, x):
    """The scaled intersection volume of two spheres, alpha(d, x) with x = r / 2R (core/etc.cpp:76-101)."""
    if x > 1.0 or x < 0:
        return 0.0
    if d == 1:
        return 1.0 - x
    if d == 2:
        return 2.0 / math.pi * (math.acos(x) - x * math.sqrt(1.0 - x * x))
    if d == 3:
        return 1.0 - 1.5 * x + 0.5 * x * x * x
    if d == 4:
        return 2.0 / math.pi * (math.acos(x) - (5.0 * x - 2.0 * x * x * x) / 3.0 * math.sqrt(1 - x * x))
    if d == 5:
        return 1.0 - 1.875 * x + 1.25 * x * x * x - 0.875 * x * x * x * x * x
    return 0.0


class _Tokens:
    """`ifile >> x` over a text stream: whitespace-separated tokens."""

    def __init__(self, stream):
        text = stream.read() if hasattr(stream, "read") else str(stream)
        self._t = text.split()
        self._i = 0

    def next(self):
        if self._i >= len(self._t):
            raise EOFError("GetCCO: the input ended early")
        v = self._t[self._i]
        self._i += 1
        return v


def GetCCO(stream, seed=0, timelimit_in_hour=144, beg_idx=0, vtilde="flat", out=None):
    """Run GetCCO on the input `stream` (the text the reference reads from standard input).  seed, beg_idx and
    vtilde are the reference's command-line arguments 2, 3 and 5 (main.cpp:145-160).  Returns the result of the
    run: the MultiRun dictionary (ground) or the list of (potential, kinetic) energies of the samples (MD)."""
    from .formats import ConfigurationPack
    from .md import CollectiveCoordinateMD
    from .optimize import CollectiveCoordinateMultiRun
    out = out or sys.stdout
    tok = _Tokens(stream)
    rngGod = RandomGenerator(seed)
    dim = int(tok.next())
    K1 = float(tok.next())
    K2 = float(tok.next())
    S0 = float(tok.next())
    val = float(tok.next())
    sigma = float(tok.next())
    phi = float(tok.next())
    tok.next()  # number of threads: the engine runs on the GPU
    num = int(tok.next())
    numConfig = int(tok.next())
    L = math.pow(num, 1.0 / dim)
    a = math.pow(phi / HyperSphere_Volume(dim, 1.0), 1.0 / dim)
    init = tok.next()
    load = None
    if init == "random":
        def GetInitConfigs(i):
            return RandomConfiguration(rngGod, num, dim, L)
    elif init == "input":
        load = ConfigurationPack(tok.next())

        def GetInitConfigs(i):
            return load.GetConfig(i)
    else:
        raise ValueError(f"{init} is undefined")

    def with_cells(c):
        # cells of max(sigma, (V/N)^(1/d)) for the contact search (main.cpp:234-243)
        n = max(1.0, float(c.NumParticle()))
        c.cell_size = max(sigma, (c.PeriodicVolume() / n) ** (1.0 / c.GetDimension()))
        return c

    savename = tok.next()
    mode = tok.next()
    print(f"NumParticle = {num}\ndirectory = {savename}\na = {a}", file=out)

    # the potential (main.cpp:250-318): constraints from GetKs of the first start with K1/a < |k| <= K2/a
    potential = RepulsiveCCPotential(dim, val, sigma, S0)
    k1, k2 = K1 / a, K2 / a
    config0 = with_cells(GetInitConfigs(beg_idx))
    if S0 == 0.0:
        if vtilde == "flat":
            v = lambda k: 1.0  # noqa: E731
        elif vtilde == "overlap":
            v = lambda k: alpha(dim, k / k2)  # noqa: E731
        elif vtilde == "power-law":
            m = float(tok.next())
            v = lambda k: math.pow(1.0 - k / k2, m)  # noqa: E731
        else:
            raise ValueError(f"unknown vtilde function '{vtilde}'")
    chi = 0
    for kk in GetKs(config0, k2, k2, 1):
        if float(kk @ kk) > k1 * k1:
            knorm = math.sqrt(float(kk @ kk))
            potential.CCPotential.AddConstraint(kk, [v(knorm)] if S0 == 0.0 else [1.0, S0])
            chi += 1
    chi /= dim * (num - 1)
    print(f"chi = {chi}", file=out)

    if mode == "MD":
        dt, T, steps, equil, auto = 0.01, -1.0, 100000, 500, False
        while True:
            cmd = tok.next().lower()
            if cmd == "run":
                break
            if cmd == "timestep":
                dt = float(tok.next())
            elif cmd == "samplesteps":
                steps = int(tok.next())
            elif cmd == "temperature":
                T = float(tok.next())
            elif cmd == "numequilibration":
                equil = int(tok.next())
            elif cmd == "autotimestep":
                auto = True
            elif cmd == "restore":
                print("the MDDump restart file is not supported; running from the start", file=out)
            elif cmd == "finufft":
                potential.SetRhoBackend("nufft")
            else:
                print(f"{cmd} is an undefined command", file=out)
        if T < 0.0:
            T = {1: 2e-4, 2: 2e-6, 3: 1e-6}.get(dim, 1e-6)
        if init == "random":
            # a fresh start from the generator, as the reference's GetInitConfigs(0) (main.cpp:384-390)
            start = with_cells(GetInitConfigs(0))
            return CollectiveCoordinateMD(start, potential, rngGod, dt, T, savename, numConfig, steps, equil, auto)
        start = with_cells(GetInitConfigs(load.NumConfig() - 1))
        done = ConfigurationPack(savename).NumConfig()
        return CollectiveCoordinateMD(start, potential, rngGod, dt, T, savename, max(numConfig - done, 0), steps,
                                      equil, auto)
    if mode == "ground":
        max_steps, tolerance, algorithm = 10000, 1e-14, "LBFGS"
        while True:
            cmd = tok.next().lower()
            if cmd == "run":
                break
            if cmd == "tolerance":
                tolerance = float(tok.next())
            elif cmd == "maxsteps":
                max_steps = int(tok.next())
            elif cmd == "algorithm":
                algorithm = tok.next()
            elif cmd == "finufft":
                potential.SetRhoBackend("nufft")
            else:
                print(f"{cmd} is an undefined command", file=out)
        print(f"Numerical optimizer is {algorithm}\nEnergy tolerance is {tolerance}\n"
              f"Max. steps of evaluations is {max_steps}", file=out)
        if init == "random":
            have = ConfigurationPack(savename + "_Success").NumConfig()
            todo = max(numConfig - have, 0)
            print(f"We already have {have} Configurations in the savefile\njust compute {todo} more Configurations",
                  file=out)
            start = with_cells(GetInitConfigs(0))  # GetInitConfigs(0) draws a new start (main.cpp:486)
            return CollectiveCoordinateMultiRun(start, potential, rngGod, savename, todo, None, algorithm,
                                                tolerance, max_steps)
        used = ConfigurationPack(savename + "_InitConfig").NumConfig()
        have = ConfigurationPack(savename + "_InitConfig").NumConfig()
        results = []
        ident = used + beg_idx
        for _ in range(have, numConfig):
            if ident < load.NumConfig():
                c = with_cells(GetInitConfigs(ident))
                results.append(CollectiveCoordinateMultiRun(c, potential, rngGod, savename, 1, None, algorithm,
                                                            tolerance, max_steps))
            ident += 1
        return results
    raise ValueError(f"mode {mode} is undefined!")

