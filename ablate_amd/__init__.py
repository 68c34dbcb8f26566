"""ablate_amd: MI355X-native finite-volume compressible-flow RHS (drop-in for ABLATE's
FiniteVolumeSolver::ComputeRHSFunction), HIP kernels for gfx950 behind the C ABI in include/ablate_fv.h."""
from . import _lib
from .fv import (
    Ausm,
    AverageFlux,
    Constant,
    FiniteVolumeSolver,
    InflowGhost,
    NavierStokesTransport,
    OffFlux,
    OutflowGhost,
    PerfectGas,
    Rusanov,
    SpeciesTransport,
    Sutherland,
)
from .mesh import BoxMesh, TetBoxMesh

__all__ = [
    "Ausm", "AverageFlux", "BoxMesh", "Constant", "FiniteVolumeSolver", "InflowGhost", "NavierStokesTransport",
    "OffFlux", "OutflowGhost", "PerfectGas", "Rusanov", "SpeciesTransport", "Sutherland", "TetBoxMesh",
]
