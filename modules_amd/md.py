"""Molecular dynamics over the CC potentials: ParticleMolecularDynamics (potential/MD_System.h:23-131) with its
state resident on the GPU, and the CollectiveCoordinateMD driver (apps/CCOptimization/main.cpp:1875-2010).

Every step (velocity Verlet, Nose-Hoover, Anderson) runs on the device through the library's MD handle
(include/ccgpu.h); random draws come from the caller's RandomGenerator in the reference's order.
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from ._lib import check, lib
from .potential import Configuration, _p


class ParticleMolecularDynamics:
    """ParticleMolecularDynamics(Position, TimeStep, Mass): a copy of the configuration, zero velocities and
    accelerations, one mass for every particle (MD_System.cpp:17-29)."""

    def __init__(self, Position: Configuration, TimeStep, Mass=1.0):
        self.Dimension = Position.GetDimension()
        self.NumParticle = Position.NumParticle()
        self._basis = np.ascontiguousarray(Position.basis, dtype=np.float64).copy()
        # the Position's cell-list size and particle names travel with it into the restart file
        self._cell_size = getattr(Position, "cell_size", None)
        self._names = getattr(Position, "names", None)
        m = np.asarray(Mass, dtype=np.float64)
        self.Mass = np.full(self.NumParticle, float(m)) if m.ndim == 0 else m.copy()
        rel = np.ascontiguousarray(Position.relative, dtype=np.float64)
        self._h = C.c_void_p()
        check(lib().ccgpu_md_create(self.Dimension, _p(self._basis), self.NumParticle, _p(rel), float(TimeStep),
                                    _p(np.ascontiguousarray(self.Mass)), C.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            from . import _lib
            if _lib._lib is not None:
                lib().ccgpu_md_destroy(h)
            self._h = None

    # --- state -----------------------------------------------------------------------------------
    def _state(self):
        n, d = self.NumParticle, self.Dimension
        rel, v, a = np.zeros((max(n, 1), d)), np.zeros((max(n, 1), d)), np.zeros((max(n, 1), d))
        t, dt, xi = C.c_double(0), C.c_double(0), C.c_double(0)
        check(lib().ccgpu_md_get_state(self._h, _p(rel), _p(v), _p(a), C.byref(t), C.byref(dt), C.byref(xi)))
        return rel[:n], v[:n], a[:n], t.value, dt.value, xi.value

    @property
    def Position(self):
        c = Configuration(self._basis, self._state()[0])
        if self._cell_size:
            c.cell_size = self._cell_size
        if self._names is not None:
            c.names = self._names
        return c

    @property
    def Velocity(self):
        return self._state()[1]

    @Velocity.setter
    def Velocity(self, v):
        check(lib().ccgpu_md_set_velocities(self._h, _p(np.ascontiguousarray(v, dtype=np.float64).reshape(-1))))

    @property
    def Acceleration(self):
        return self._state()[2]

    @property
    def Time(self):
        return self._state()[3]

    @property
    def TimeStep(self):
        return self._state()[4]

    @TimeStep.setter
    def TimeStep(self, dt):
        check(lib().ccgpu_md_set_timestep(self._h, float(dt)))

    @property
    def xi(self):
        return self._state()[5]

    @property
    def MassQ(self):
        q = C.c_double(0)
        check(lib().ccgpu_md_get_q(self._h, C.byref(q)))
        return q.value

    # --- restart file (Prefix.MDDump) ------------------------------------------------------------
    def WriteDump(self, path, stage, step):
        """The restart file of CollectiveCoordinateMD: stage and step, then WriteBinary (potential/MD_System.cpp:50-62;
        layout in include/ccgpu.h, ccgpu_mddump_write)."""
        from .formats import ConfigurationPack, _default_cell_size
        rel, v, a, t, dt, xi = self._state()
        pos = self.Position
        chars = ConfigurationPack()._chars(pos)
        n = self.NumParticle
        check(lib().ccgpu_mddump_write(os.fsencode(str(path)), int(stage), int(step), self.Dimension,
                                       _default_cell_size(pos), _p(self._basis), n, _p(np.ascontiguousarray(rel)),
                                       None if chars is None else chars.ctypes.data_as(C.c_void_p), dt, t, xi,
                                       self.MassQ, _p(np.ascontiguousarray(self.Mass)), _p(np.ascontiguousarray(v)),
                                       _p(np.ascontiguousarray(a))))

    @classmethod
    def ReadDump(cls, path):
        """(stage, step, system) from a restart file: ParticleMolecularDynamics(std::istream &)
        (potential/MD_System.cpp:34-49) after the stage and step (apps/CCOptimization/main.cpp:1893-1899)."""
        b = os.fsencode(str(path))
        st, sp, d, n = C.c_int32(0), C.c_int64(0), C.c_int(0), C.c_int64(0)
        cs, dt, t, xi, q = C.c_double(0), C.c_double(0), C.c_double(0), C.c_double(0), C.c_double(0)
        basis = np.zeros(16)
        args = lambda cap, rel, chars, m, v, a: (b, C.byref(st), C.byref(sp), C.byref(d), C.byref(cs), _p(basis),
                                                 C.byref(n), cap, rel, chars, C.byref(dt), C.byref(t), C.byref(xi),
                                                 C.byref(q), m, v, a)
        check(lib().ccgpu_mddump_read(*args(0, None, None, None, None, None)))
        N, dd = n.value, d.value
        rel, v, a = (np.zeros((max(N, 1), dd)) for _ in range(3))
        mass = np.zeros(max(N, 1))
        chars = np.zeros(max(N, 1) * 4, dtype=np.uint8)
        check(lib().ccgpu_mddump_read(*args(N, _p(rel), chars.ctypes.data_as(C.c_void_p), _p(mass), _p(v), _p(a))))
        pos = Configuration(basis[: dd * dd].reshape(dd, dd), rel[:N])
        pos.cell_size = cs.value
        pos.names = [bytes(chars[4 * i: 4 * i + 4]).split(b"\0")[0] for i in range(N)]
        system = cls(pos, dt.value, mass[:N])
        check(lib().ccgpu_md_restore(system._h, _p(np.ascontiguousarray(v[:N])), _p(np.ascontiguousarray(a[:N])),
                                     t.value, xi.value, q.value))
        return st.value, sp.value, system

    # --- dynamics --------------------------------------------------------------------------------
    def Evolve(self, repeat, potential):
        potential._sync()
        check(lib().ccgpu_md_evolve(self._h, potential._h, int(repeat)))
        self._rebind(potential)

    def InitNoseHoover(self, Q):
        check(lib().ccgpu_md_init_nose_hoover(self._h, float(Q)))

    def NoseHooverEvolve(self, repeat, potential, kT):
        potential._sync()
        check(lib().ccgpu_md_nose_hoover_evolve(self._h, potential._h, int(repeat), float(kT)))
        self._rebind(potential)

    def AndersonEvolve(self, repeat, potential, kT, ResetSpeedProbability, gen):
        """Resets draw from `gen` in the reference's order; the generator advances by what the resets used."""
        count, resets = 0.0, 0
        for _ in range(int(repeat)):
            count += ResetSpeedProbability
            while count > 0:
                resets += 1
                count -= 1.0
        n = resets * (1 + 2 * self.Dimension) + 64  # a draw of exactly 0 is redrawn (rare): keep some spare
        gen._ensure(n)
        draws = np.ascontiguousarray(gen._buf[gen._pos:gen._pos + n])
        used = C.c_int64(0)
        potential._sync()
        check(lib().ccgpu_md_anderson_evolve(self._h, potential._h, int(repeat), float(kT),
                                             float(ResetSpeedProbability), _p(draws), n, C.byref(used)))
        gen._pos += used.value
        self._rebind(potential)

    def _rebind(self, potential):
        # the potential is bound to the final positions (potential.SetConfiguration(Position) of the last step)
        potential._config = self.Position
        potential.RhoValid = False
        if hasattr(potential, "_pull"):  # variable-cell potentials mirror their regenerated constraints
            potential._pull()

    def GetKineticEnergy(self):
        e = C.c_double(0)
        check(lib().ccgpu_md_kinetic_energy(self._h, C.byref(e)))
        return e.value

    def GetPotentialEnergy(self, potential):
        potential.SetConfiguration(self.Position)
        return potential.Energy()

    def SetRandomSpeed(self, kT, gen):
        n = 2 * self.NumParticle * self.Dimension + 64  # |v| > 1e30 redraws (a zero draw) are rare
        gen._ensure(n)
        draws = np.ascontiguousarray(gen._buf[gen._pos:gen._pos + n])
        used = C.c_int64(0)
        check(lib().ccgpu_md_set_random_speed(self._h, float(kT), _p(draws), n, C.byref(used)))
        gen._pos += used.value

    def RemoveOverallMomentum(self):
        check(lib().ccgpu_md_remove_overall_momentum(self._h))

    def Evolve_AutoTimeStep(self, repeat, potential, DeltaLogEnergyPerStep):
        """Evolve_AutoTimeStep (MD_System.cpp:172-201): the step follows the drift of the total energy."""
        Ebefore = self.GetKineticEnergy() + self.GetPotentialEnergy(potential)
        self.Evolve(repeat, potential)
        Eafter = self.GetKineticEnergy() + self.GetPotentialEnergy(potential)
        dlog = abs(np.log(Ebefore / Eafter) / repeat)
        dt = self.TimeStep
        if dlog > 5 * DeltaLogEnergyPerStep:
            dt *= 0.9
        elif dlog > 3 * DeltaLogEnergyPerStep:
            dt *= 0.95
        elif dlog > 2 * DeltaLogEnergyPerStep:
            dt *= 0.99
        elif dlog < 0.01 * DeltaLogEnergyPerStep:
            dt *= 1.05
        elif dlog < 0.5 * DeltaLogEnergyPerStep:
            dt *= 1.005
        self.TimeStep = dt


def CollectiveCoordinateMD(pConfig: Configuration, pPotential, gen, TimeStep, Temperature, Prefix, SampleNumber,
                           StepPerSample, EquilibrateSamples, MDAutoTimeStep=False, AllowRestore=False, TimeLimit=None):
    """CollectiveCoordinateMD (apps/CCOptimization/main.cpp:1875-2012): relax to a ground state
    (RelaxStructure_NLOPT, 1000 evaluations per pass), equilibrate EquilibrateSamples times (two rounds of
    SetRandomSpeed + Evolve(StepPerSample / 2)), then sample SampleNumber configurations (AndersonEvolve at 0.01 reset
    probability + Evolve, each StepPerSample / 2 steps) into the ConfigurationPack `Prefix`.

    AllowRestore: continue from Prefix.MDDump when it exists (its stage and the step after its step; the pack is
    then kept, not cleared), and store the final state there at the end.  TimeLimit (seconds since the epoch, as
    the reference's time_t; None: none): once passed, the state is stored in Prefix.MDDump with the stage and the
    step about to run, "Not Completed. Please run again" goes to Prefix_continue.txt (equilibration) or
    continue.txt (sampling), and the function returns -- the reference's behaviour, file names included.  The
    generator's state is not part of the file, as in the reference.  Returns the list of (potential energy,
    kinetic energy) of the samples taken in this call."""
    from .formats import ConfigurationPack
    from .optimize import RelaxStructure_NLOPT
    pack = ConfigurationPack(Prefix)
    dump = str(Prefix) + ".MDDump"
    system, stage, step = None, 0, 0
    if AllowRestore and os.path.isfile(dump):
        stage, step, system = ParticleMolecularDynamics.ReadDump(dump)
        pPotential.SetConfiguration(system.Position)
        step += 1
    if system is None:
        pack.Clear()
        RelaxStructure_NLOPT(pConfig, pPotential, 0.0, 0, 0.0, 1000)
        pPotential.SetConfiguration(pConfig)
        pPotential.Energy()
        system = ParticleMolecularDynamics(pConfig, TimeStep, 1.0)
    expired = lambda: TimeLimit is not None and time.time() > TimeLimit
    half = StepPerSample // 2
    if stage == 0:
        stage, step = 1, 0
    out = []
    if stage == 1:
        for i in range(step, EquilibrateSamples):
            if expired():
                system.WriteDump(dump, stage, i)
                with open(str(Prefix) + "_continue.txt", "w") as f:
                    f.write("Not Completed. Please run again\n")
                return out
            for _ in range(2):
                system.SetRandomSpeed(Temperature, gen)
                if MDAutoTimeStep:
                    system.Evolve_AutoTimeStep(half, pPotential, 0.0001)
                else:
                    system.Evolve(half, pPotential)
        stage, step = 2, 0
    final_step = 0
    for i in range(step, SampleNumber):
        if expired():
            system.WriteDump(dump, stage, i)
            with open("continue.txt", "w") as f:
                f.write("Not Completed. Please run again\n")
            return out
        system.AndersonEvolve(half, pPotential, Temperature, 0.01, gen)
        system.Evolve(half, pPotential)
        result = system.Position
        pPotential.SetConfiguration(result)
        out.append((pPotential.Energy(), system.GetKineticEnergy()))
        pack.AddConfig(result)
        final_step = i
    if AllowRestore:
        system.WriteDump(dump, stage, final_step)
    return out
