"""sdc_amd: MI355X-native Rayleigh-Benard time stepper (drop-in for gpu_flow's hot path).

The compute path is libsdc_rbc.so (hand-written HIP for gfx950 behind the C-ABI in
include/rbc_abi.h); this package holds its sources (csrc/), the build script and the
Python host binding.
"""
from .rbc import (Closure, Ensemble, Stepper, Tracer, RbcError, RbcNumericError, load_library,  # noqa: F401
                  LIB_PATH)

__all__ = ["Closure", "Ensemble", "Stepper", "Tracer", "RbcError", "RbcNumericError", "load_library", "LIB_PATH"]
