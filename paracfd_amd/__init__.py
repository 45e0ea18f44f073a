"""paracfd_amd: MI355X-native solver step for ParaCFD's external-aerodynamics path.

The compute path is libparacfd.so (HIP kernels for gfx950 behind the C ABI of
include/paracfd.h).  Import never falls back to a CPU implementation: the library must be
built (see __graft_entry__.build()).
"""
from . import _lib, cases
from ._lib import PcfdError, Topology
from .core import ExternalAeroCore, default_options

__all__ = ["ExternalAeroCore", "Topology", "PcfdError", "default_options", "cases", "_lib"]
