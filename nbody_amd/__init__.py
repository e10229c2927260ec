"""nbody_amd -- MI355X (gfx950) FMM force pipeline behind the nbody Simulation
interface.  The compute path is libnbody_hip.so (HIP kernels + C ABI in
include/nbody_hip.h); this package is its Python harness."""
from ._lib import (Config, Stats, default_config, lib, SOLVER_FMM, SOLVER_DIRECT, M2L_PLAIN, M2L_GROUPED,
                   M2L_BF16X3, M2L_MFMA_F32, M2L_TERMS_ALL, M2L_TERMS_LOW, P2P_LEAF, P2P_LANE, INT_KDK, INT_EULER, INT_KICK_DRIFT, LAW_SOFTENED, LAW_COMPAT_NAIVE)
from .simulation import HipSimulation, Particle, Simulation, device_count
from . import ics
from .cpu import CpuSimulation

__all__ = ["Config", "Stats", "default_config", "lib", "HipSimulation", "CpuSimulation", "Particle", "Simulation",
           "device_count",
           "ics", "SOLVER_FMM", "SOLVER_DIRECT", "M2L_PLAIN", "M2L_GROUPED", "M2L_BF16X3", "M2L_MFMA_F32", "M2L_TERMS_ALL", "M2L_TERMS_LOW", "P2P_LEAF", "P2P_LANE", "INT_KDK", "INT_EULER",
           "INT_KICK_DRIFT", "LAW_SOFTENED", "LAW_COMPAT_NAIVE"]
