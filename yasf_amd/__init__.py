"""yasf_amd: MI355X-native hot path of the multiple-sphere T-matrix solve.

A drop-in ``hip`` coupling backend plus device GMRES and block-diagonal
preconditioner behind a mirror of the reference's Simulation/Solver API
(AGBV/YASF ``yasfpy``).  Device work runs in ``libyasfhip.so`` (hand-written
HIP for gfx950) through the C ABI in ``include/yasfhip.h``.
"""

from yasf_amd.model import InitialField, Parameters, Particles
from yasf_amd.numerics import Numerics
from yasf_amd.preconditioner import BlockDiagonalPreconditioner, edge_size_for, make_particle_partition
from yasf_amd.simulation import Simulation
from yasf_amd.solver import Solver

__all__ = [
    "BlockDiagonalPreconditioner", "InitialField", "Numerics", "Parameters", "Particles",
    "Simulation", "Solver", "edge_size_for", "make_particle_partition",
]
