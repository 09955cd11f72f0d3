"""MI355X-native SWEHR coupled time step.

The product is the C ABI in include/swehr.h, implemented by swehr_amd/libswehr_hip.so
(hand-written HIP kernels for gfx950), and the drop-in program swehr_amd/bin/swehr.
This package adds Python access for tests and benchmarks (``solver``), the
reference's file contract and seeded synthetic cases (``inputs``).
"""
from . import _abi  # noqa: F401
from .inputs import Case, read_case, write_case  # noqa: F401

__all__ = ["Case", "read_case", "write_case"]
