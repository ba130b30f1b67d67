"""cusna_amd -- MI355X (gfx950) SAOM period simulator behind the libcusna C ABI.

Layout:
  csrc/        HIP kernels (period simulator, statistics) and the C ABI host code
  build.py     hipcc build of _lib/libcusna_amd.so for gfx950
  lib.py       ctypes bindings of the library (no CPU fallback)
  abi.py       ctypes mirror of the ABI structs
  data.py      SAOM panel data conventions (restates the reference's R/data.R)
  effects.py   effect registry and descriptors (restates R/effects.R)
  period.py    one period call from data + effects (restates R/mom.R:91-125)
"""
__version__ = "0.1.0"
