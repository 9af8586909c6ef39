"""MI355X-native bmad_standard multi-particle FP64 tracking.

Host side of the drop-in for Bmad's beam tracking path (track_beam / track_bunch /
track1_bunch / track1, bmad/multiparticle/beam_mod.f90 and bmad/code/track1.f90): element
records and bookkeeping (``elements``), the benchmark rings and beams (``lattices``), the
ctypes binding of the native library libbmx.so (``native``) and the host API (``tracking``).
"""
from . import elements, flops, lattices, native, serving, tracking  # noqa: F401

__all__ = ["elements", "flops", "lattices", "native", "tracking"]
