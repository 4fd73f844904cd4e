"""Host side of the MI355X RK3 (magneto)hydrodynamics timestep.

The compute path is the HIP library ``libmhd3d.so`` (csrc/, C ABI in
include/mhd3d.h); this package builds run descriptors, initial states and
drives the library.
"""
from .desc import RunConfig, Descriptor, MhdDesc, default_par, set_constants, f32, fdiv  # noqa: F401
