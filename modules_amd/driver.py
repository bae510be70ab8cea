"""GetCCO and GetMultiCCO, the drivers of apps/CCOptimization/main.cpp:135-620 and :1052-1707, reading the same
input streams.

The reference reads its parameters from standard input token by token (main.cpp:182-249) and then the commands of
the chosen mode until `run` (MD: :345-404, ground: :452-477).  This driver reads the same tokens from a text stream,
builds the same RepulsiveCCPotential (constraints K1/a < |k| < K2/a weighted by the flat, overlap or power-law
tildeV, or the target-S(k) form when S0 != 0, :250-318) and runs CollectiveCoordinateMultiRun (ground) or
CollectiveCoordinateMD (MD) on the engine, writing the reference's ConfigurationPacks under the save name.
"""
from __future__ import annotations

import math
import re
import sys
import time

import numpy as np

from .potential import Configuration, GetKs, RandomConfiguration, RandomGenerator, RepulsiveCCPotential

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
    # main.cpp:179: the run's time limit, five minutes short of timelimit_in_hour
    TimeLimit = time.time() + timelimit_in_hour * 3600 - 5 * 60
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
        dt, T, steps, equil, auto, restore = 0.01, -1.0, 100000, 500, False, False
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
                restore = True
            elif cmd == "finufft":
                potential.SetRhoBackend("nufft")
            else:
                print(f"{cmd} is an undefined command", file=out)
        if T < 0.0:
            T = {1: 2e-4, 2: 2e-6, 3: 1e-6}.get(dim, 1e-6)
        if init == "random":
            # a fresh start from the generator, as the reference's GetInitConfigs(0) (main.cpp:384-390)
            start = with_cells(GetInitConfigs(0))
            return CollectiveCoordinateMD(start, potential, rngGod, dt, T, savename, numConfig, steps, equil, auto,
                                          AllowRestore=restore, TimeLimit=TimeLimit)
        start = with_cells(GetInitConfigs(load.NumConfig() - 1))
        done = ConfigurationPack(savename).NumConfig()
        return CollectiveCoordinateMD(start, potential, rngGod, dt, T, savename, max(numConfig - done, 0), steps,
                                      equil, auto, AllowRestore=restore, TimeLimit=TimeLimit)
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



def _atoi(s):
    """atoi: the leading decimal integer of s (0 when there is none)."""
    m = re.match(r"\s*([+-]?\d+)", s)
    return int(m.group(1)) if m else 0


def _atof(s):
    """atof: the leading decimal number of s (0 when there is none)."""
    m = re.match(r"\s*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)", s)
    return float(m.group(1)) if m else 0.0


def GetMultiCCO(stream, timelimit_in_hour=144, seed=0, beg_idx=0, verbosity=0, vtilde="flat", out=None):
    """Run GetMultiCCO (apps/CCOptimization/main.cpp:1052-1707), the multi-species driver, on the input `stream`.
    timelimit_in_hour, seed, beg_idx, verbosity and vtilde are its command-line arguments 2-6 (:1061-1077).

    The stream holds, as the reference reads it: the initial packing ("random": dimension, the number of radii,
    then number and relative radius of each, then the packing fraction, :1101-1146; "input": a SpherePacking text
    name with %d for the index, :1148-1168), the contact strength, the number of species and, per species, its
    particle count, S(0), "chi" or "wavenumber", lower and upper bounds (and the exponent of a power-law tildeV),
    :1181-1286; then # threads, the save name, the number of configurations and the mode with its commands.

    Ground mode relaxes with CollectiveCoordinateMultiRun over the MultiStealthyPotential on the batched device
    solver, exports accepted configurations (export txt|binary|pack) and splits the success pack by species
    (savename_s<s>_Success) unless `combined` is given (:1420-1610).  Returns (status, result): status is the
    reference's return code (1 where it prints an error and returns 1) and result the MultiRun dictionary or None.

    As in the reference:
      * the MD branch chooses its start by the last token read into its scratch string (the last species' "chi" /
        "wavenumber" word, :1384, :1401), not by the initial-packing word, so it runs nothing unless that word is
        "random" or "input";
      * the "input" packing never sets the count of loadable configurations, so both modes stop with the error
        the reference prints (:1398-1400, :1620-1623).
    A "random" word there would run molecular dynamics over the mixture, which this engine does not integrate
    (its MD handle takes single-configuration potentials): that case raises NotImplementedError.  `finufft` /
    `finufft-sequential` / `paired-forces` keep the direct sums (the exact values the reference's type-1 transforms
    approximate within their tolerance) and `pairstat` raises NotImplementedError: the pair-statistics report
    belongs to the analysis tools outside this engine.
    """
    from . import optimize
    from .formats import ConfigurationPack, MSPB_Pack, MSPBSnapshot, ReadSpherePacking
    from .multi import MultiStealthyPotential, SpherePacking
    from .optimize import CollectiveCoordinateMultiRun
    from .potential import getK
    out = out or sys.stdout
    tok = _Tokens(stream)
    rngGod = RandomGenerator(seed)
    print(f"Time limit for simulations is {timelimit_in_hour} hours\nRandom seed is {seed}\n"
          f"Verbosity is {verbosity}\nShape of potential is {vtilde}", file=out)
    TimeLimit = time.time() + timelimit_in_hour * 3600 - 5 * 60
    scratch = ""  # the reference's char tempstring[1000]

    # 1. the initial packing (main.cpp:1098-1175)
    init = tok.next()
    if init == "random":
        dim = int(tok.next())
        num_radii = int(tok.next())
        sizes, radii = [], []
        for _ in range(num_radii):
            scratch = tok.next()
            sizes.append(_atoi(scratch))
            scratch = tok.next()
            radii.append(_atof(scratch))
        N = sum(sizes)
        phi = float(tok.next())
        curr_phi = sum(n * HyperSphere_Volume(dim, r) for n, r in zip(sizes, radii)) / N
        factor = math.pow(phi / curr_phi, 1.0 / dim)
        radii = [r * factor for r in radii]
        print("-- Particle number and radius --", file=out)
        for n, r in zip(sizes, radii):
            print(f"{n}\t{r}", file=out)
        L = math.pow(N, 1.0 / dim)

        def GetInitConfigs(i):
            # SpherePacking(GetUnitCubicBox(dim, 0.1)).Rescale(L), then each radius class in order, d draws per
            # sphere (RandomVectorInCube)
            rel, rad = [], []
            for n, r in zip(sizes, radii):
                if n > 0:
                    rel.append(rngGod.RandomDoubles(n * dim).reshape(n, dim))
                    rad += [r] * n
            rel = np.concatenate(rel) if rel else np.zeros((0, dim))
            pck = SpherePacking(np.eye(dim) * L, rel, np.array(rad))
            pck.cell_size = 0.1 * L
            return pck
    elif init == "input":
        loadname = tok.next()

        def GetInitConfigs(idx):
            return ReadSpherePacking(loadname.replace("%d", str(idx), 1))
    else:
        print(f"{init} is undefined ", file=out)
        return 1, None
    val = float(tok.next())

    # 2. the potential (main.cpp:1178-1286)
    num_species = int(tok.next())
    c = GetInitConfigs(beg_idx)
    dim = c.GetDimension()
    N = c.NumParticle()
    rho = N / c.PeriodicVolume()
    potential = MultiStealthyPotential(dim, val)
    potential.SetPacking(c)
    potential.SetNumSpecies(num_species)
    N_curr = 0
    for s in range(num_species):
        N_s = int(tok.next())
        if N_s + N_curr > N:
            print("The input is larger than the unspecified particle number. All unspecified particles are taken to "
                  "be a new species.", file=out)
            N_s = N - N_curr
        for i in range(N_s):
            potential.SetSpeciesOfParticle(N_curr + i, s)
        N_curr += N_s
        S0 = float(tok.next())
        scratch = tok.next()
        lower = float(tok.next())
        upper = float(tok.next())
        if scratch == "chi":
            # chi is normalized by the total number of particles here
            lower = getK(lower, rho, dim)
            upper = getK(upper, rho, dim)
        if S0 == 0.0:
            if vtilde == "flat":
                v = lambda k: [1.0]  # noqa: E731
            elif vtilde == "overlap":
                v = lambda k, u=upper: [alpha(dim, k / u)]  # noqa: E731
            elif vtilde == "power-law":
                m = float(tok.next())
                v = lambda k, u=upper, m=m: [math.pow(1.0 - k / u, m)]  # noqa: E731
            else:
                print(f"Error: unknown vtilde function '{vtilde}'.", file=out)
                return 1, None
        else:
            # the species potentials are plain ShiftedCCPotentials: S0 rides along and the weight is 1
            v = lambda k: [1.0, S0]  # noqa: E731
        chi = 0.0
        for kk in GetKs(c, upper, upper, 1):
            m2 = float(kk @ kk)
            if m2 > lower * lower:
                potential.AddConstraint(s, kk, v(math.sqrt(m2)))
                chi += 1
        chi = chi / (dim * (N_s - 1)) if N_s >= 2 else 0.0
        print(f"chi (of a subconfig) = {chi}", file=out)

    tok.next()  # number of threads: the engine runs on the GPU
    savename = tok.next()
    N_config = int(tok.next())
    mode = tok.next()
    print(f"directory = {savename}", file=out)

    if mode == "MD":
        T = -1.0
        while True:
            cmd = tok.next().lower()
            if cmd == "run":
                break
            if cmd in ("timestep", "samplesteps", "temperature", "numequilibration"):
                value = tok.next()
                if cmd == "temperature":
                    T = float(value)
            elif cmd in ("autotimestep", "restore", "finufft", "finufft-sequential", "paired-forces"):
                pass
            else:
                print(f"{cmd} is an undefined command", file=out)
        if T < 0.0:
            T = {1: 2e-4, 2: 2e-6, 3: 1e-6}.get(dim, T)
        print(f"Input T = {T}", file=out)
        if scratch == "random":
            raise NotImplementedError("GetMultiCCO: molecular dynamics over a multi-species potential")
        if scratch == "input":
            print("Error: input mode selected but the resolved input configuration count (N_load) is 0; cannot "
                  "continue from a previous run.", file=out)
            return 1, None
        return 0, None

    if mode != "ground":
        print(f"mode {mode} is undefined!", file=out)
        return 1, None
    max_steps, tolerance, algorithm = 10000, 1e-14, "LBFGS"
    split_species = True
    export_enabled = export_pack = False
    export_format = "text"
    archive = MSPB_Pack()
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
        elif cmd == "combined":
            split_species = False
        elif cmd in ("finufft", "finufft-sequential", "paired-forces"):
            print("Collective densities keep the direct sums.", file=out)
        elif cmd == "lbfgs":
            name = tok.next()
            codes = {"nlopt": 0, "ring": 1, "fast": 3}
            if name not in codes:
                print("Expected: lbfgs nlopt|ring|fast", file=out)
                return 1, None
            optimize.StructureOptimization_LBFGSImplementation = codes[name]
        elif cmd == "history":
            t = tok.next()
            if not t.isdigit() or int(t) == 0:
                print("Expected: history N (N >= 1)", file=out)
                return 1, None
            optimize.StructureOptimization_LBFGSHistory = int(t)
        elif cmd == "export":
            fmt = tok.next()
            if fmt not in ("txt", "binary", "pack", "off"):
                print("Expected: export txt|binary|pack|off", file=out)
                return 1, None
            export_enabled = fmt != "off"
            export_pack = fmt == "pack"
            export_format = "binary" if fmt == "binary" else "text"
        elif cmd == "pairstat":
            for _ in range(6):
                tok.next()
            raise NotImplementedError("GetMultiCCO: pairstat reports are outside this engine")
        else:
            print(f"{cmd} is an undefined command", file=out)
    print(f"\nNumerical optimizer is {algorithm}\nEnergy tolerance is {tolerance}\n"
          f"Max. steps of evaluations is {max_steps}\n"
          f"L-BFGS history in use = {optimize.StructureOptimization_LBFGSHistory}", file=out)

    def onAccepted(conf, index):
        # main.cpp:1431-1458: export each accepted configuration as the potential stores it
        if not export_enabled:
            return
        potential.SetConfiguration(conf)
        if export_pack:
            if archive.NumConfig() != index:
                raise RuntimeError("MSPB archive and success pack counts differ")
            archive.AddConfig(MSPBSnapshot.Read(potential.ExportConfiguration(None, "binary")))
        else:
            ext = ".mspb" if export_format == "binary" else ".txt"
            potential.ExportConfiguration(f"{savename}_Success_{index}{ext}", export_format)

    if init == "random":
        pck = GetInitConfigs(0)  # a new start from the generator (main.cpp:1567)
        potential.SetPacking(pck)
        pConfig = Configuration(pck.basis, pck.relative)
        N_success = ConfigurationPack(savename + "_Success").NumConfig()
        if export_pack:
            archive.Open(savename + "_Success")
            if archive.NumConfig() != N_success:
                raise RuntimeError("MSPB archive must contain all existing successes before resuming with export pack")
        todo = max(N_config - N_success, 0)
        print(f"We already have {N_success} Configurations in the savefile \njust compute {todo} more "
              f"Configurations \nmax_eval = {max_steps}", file=out)
        result = CollectiveCoordinateMultiRun(pConfig, potential, rngGod, savename, todo, TimeLimit, algorithm,
                                              tolerance, max_steps, onAccepted)
        if split_species:
            # per-species packs of the new successes, each species' particles in its list order (:1591-1611)
            cp = [ConfigurationPack(f"{savename}_s{s}_Success") for s in range(len(potential.species_constraints))]
            success = ConfigurationPack(savename + "_Success")
            idx = [sp.idx for sp in potential.species_constraints]
            for i in range(N_success, success.NumConfig()):
                conf = success.GetConfig(i)
                for s, pack in enumerate(cp):
                    c_spec = Configuration(conf.basis, conf.relative[idx[s]].reshape(-1, dim))
                    c_spec.cell_size = conf.cell_size
                    c_spec.names = ["a"] * len(idx[s])
                    pack.AddConfig(c_spec)
        return 0, result
    print("Error: input mode selected but the resolved input configuration count (N_load) is 0; nothing to "
          "process.", file=out)
    return 1, None
