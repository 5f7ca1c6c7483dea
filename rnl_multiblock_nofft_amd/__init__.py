"""rnl_multiblock_nofft_amd -- MI355X-native RK3 time step of LittleHarsh's
plain FFT method (channel-flow DNS), behind the reference's subroutine
interfaces.  Compute lives in liblhgpu.so (HIP, gfx950); this package is the
host-side driver.
"""
from . import configs, fieldio, initial  # noqa: F401
from ._lib import LhError, load  # noqa: F401
from .solver import Solver  # noqa: F401

__all__ = ["Solver", "configs", "initial", "fieldio", "load", "LhError"]
