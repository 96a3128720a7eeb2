"""mdacp_amd -- MI355X (gfx950) implementation of MDACP's short-range
Lennard-Jones step behind a C ABI (include/mdgpu.h, libmdgpu.so).

Python here is a thin ctypes layer (mdacp_amd/mdgpu.py) for tests, benchmarks
and the autotuner; the host side of the reference interface is C++: the MDUnit,
Variables, MeshList, PairList, ForceCalculator and observer classes of
mdacp_amd/facade, which the reference's own MDManager and projects compile
against.
"""
from .mdgpu import (FP64, MIXED, NVE, NOSEHOOVER, LANGEVIN, Domain, MDGPUError, load)  # noqa: F401
from .lattice import fcc  # noqa: F401
