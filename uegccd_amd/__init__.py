"""uegccd_amd -- MI355X (gfx950) CCD amplitude iteration for the plane-wave
uniform electron gas, a drop-in for the reference's DrvCCD.

The compute path is libuegccd.so (HIP kernels + C ABI, include/uegccd.h);
this package is its host-side mirror of the reference interface:
  uegccd_amd.inputfile  the Input format (source/IO.f90)
  uegccd_amd.driver     init_heg / DrvMBPT / DrvCCD / single_point
  uegccd_amd.lib        the ctypes binding
"""
from .inputfile import InputError, UEGInfo, parse_input, read_input, standard_input  # noqa: F401

__all__ = ["InputError", "UEGInfo", "parse_input", "read_input", "standard_input"]
