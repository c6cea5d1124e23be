"""MI355X-native OpenSWPC time-stepping core: HIP kernels behind a C ABI
(include/swpc_hip.h) with a host-side mirror of the reference's setup,
file formats and per-step module interface."""
from .medium import Grid2D, Grid3D  # noqa: F401

__version__ = "0.1.0"
