"""MI355X (gfx950) Wilson-flow and clover-observable path of LGF.

Layout:
  csrc/     HIP kernels and the C-ABI (include/lgf_flow.h) -> liblgf_flow.so
  host/     C++ host program mirroring the reference's driver API -> lgf
  lgf.py    ctypes binding used by tests and bench.py
"""
from . import lgf  # noqa: F401
