"""abt_amd: MI355X-native far-field vector diffraction for PBT (hballington12/abt).

A drop-in for the reference's `diff_mod` (src/diff_mod.f90): the HIP library
abt_amd/libabt_diff.so (C-ABI in include/abt_diff.h), its Fortran
`iso_c_binding` module abt_amd/fortran/diff_mod.f90, and this Python binding.
"""
from .beams import BeamSet, concat  # noqa: F401
from . import lib  # noqa: F401
from .lib import AbtError, DeviceGrid, diff_main, diffraction, make_far_field_bins  # noqa: F401
