"""asl_amd — MI355X-native (gfx950) LBGK hot path of ASL.

The product is ``libaslhip.so`` (hand-written HIP kernels behind the C ABI of
``include/aslhip.h``); this package is the Python host mirror of ASL's LBGK
interface over it (``lattice.py``) plus geometry helpers (``geometry.py``).
Importing works without a GPU; creating a lattice needs the built library
and a GPU and raises otherwise.
"""
from ._lib import ASLError, device_info, load_library  # noqa: F401
from .lattice import (LBGK, Batch, Block, Lattice, LBGKTurbulence, LBGKUtilities, Mlp, NumMethod,  # noqa: F401
                      VectorTemplate, d3q15, d3q19, d3q27, executeAll, generateBCConstantPressure,
                      generateBCConstantPressureVelocity, generateBCConstantVelocity, generateBCNoSlip,
                      generateBCVelocity, generateComputeSurfaceFlux, generateComputeSurfaceForce, initAll,
                      readSurface)
