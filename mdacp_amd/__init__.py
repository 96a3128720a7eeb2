"""mdacp_amd -- MI355X (gfx950) implementation of MDACP's short-range
Lennard-Jones step behind a C ABI (include/mdgpu.h, libmdgpu.so).

Python here is a thin ctypes layer for tests and benchmarks; the host side of
the reference interface is C++ (mdacp_amd/host).
"""
from .mdgpu import (FP64, MIXED, NVE, NOSEHOOVER, LANGEVIN, Domain, MDGPUError, load)  # noqa: F401
from .lattice import fcc  # noqa: F401
