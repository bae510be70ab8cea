"""modules_amd -- MI355X-native collective-coordinate stealthy-potential engine.

The hot path (rho(k), energy, forces, soft core, L-BFGS) runs in hand-written gfx950 kernels behind the
C ABI of include/ccgpu.h (modules_amd/libccgpu.so).  This package is the host-side mirror of the
reference's Potential / StructureOptimization interface over that ABI.
"""
from .potential import (Configuration, GetKs, GetUnitCubicBox, KPointConstraint, RandomConfiguration,  # noqa: F401
                        LogCCPotential, RepulsiveCCPotential, ShiftedCCPotential, ShiftedCCPotential_S2,
                        ShiftedCCPotential_S2_MultiConfig, ShiftedCCPotential_withWeight,
                        ShiftedCCPotential_ForMCMixin, get_chi, getK, pi)
from .multi import MultiStealthyPotential, SpherePacking  # noqa: F401
from .varbox import ShiftedCCPotential_varBox, VarBoxPotential  # noqa: F401
