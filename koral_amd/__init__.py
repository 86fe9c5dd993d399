"""koral_amd: MI355X-native (gfx950 HIP) implementation of KORAL's
per-timestep GRRMHD+M1 hot path behind a C-ABI (include/koral_hip.h)."""
from . import params  # noqa: F401

__version__ = "0.1.0"
