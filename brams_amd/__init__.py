"""MI355X-native CCATT gas-phase chemistry (RELACS_TUV, RODAS3) for BRAMS.

The product path is libbrams_chem.so (hand-written HIP kernels for gfx950 plus
a C-ABI, include/brams_chem.h); this package holds its sources, the code
generator, the Fortran ISO_C_BINDING shim and a thin Python mirror of the
reference call used by tests and the benchmark.
"""
from . import abi  # noqa: F401
