"""MI355X (gfx950) compute backend for Image360.

The product is the C ABI library ``image360_amd/lib/libi360hip.so``
(``include/i360hip.h``) and the C++ ``HipComputeBackend`` in
``image360_amd/host``.  This package holds the Python binding of that ABI
(``backend.HipComputeBackend``, mirroring IComputeBackend), the build step
(``build``) and the synthetic workload generators (``synth``).
"""
from ._abi import HipError, HipLibraryMissing, LIB_PATH, header_symbols, lib  # noqa: F401
from .backend import HipComputeBackend  # noqa: F401
