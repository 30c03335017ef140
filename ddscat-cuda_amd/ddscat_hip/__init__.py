"""Host-side mirror of the DDSCAT C-ABI plugin for the MI355X library.

``abi.Library`` binds the drop-in shared library (reference export set plus
``ddscat_hip_*`` extensions); ``targets`` builds the synthetic inputs the
DDSCAT Fortran driver would normally hand to that ABI.
"""
from .abi import (Library, DDSCATError, LIB_FFT3D, LIB_SLICES, REFERENCE_SYMBOLS, EXTENSION_SYMBOLS,
                  BACKEND_FFT3D, BACKEND_SLICES, PREC_MIXED, PREC_FP32, exported_symbols)
from . import targets

__all__ = ["Library", "DDSCATError", "LIB_FFT3D", "LIB_SLICES", "REFERENCE_SYMBOLS", "EXTENSION_SYMBOLS",
           "BACKEND_FFT3D", "BACKEND_SLICES", "PREC_MIXED", "PREC_FP32", "exported_symbols", "targets"]
