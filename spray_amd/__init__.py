"""spray_amd: MI355X-native per-rank tracing path of TACC/SpRay.

The compute path is libspray_gpu.so (hand-written HIP for gfx950 behind the C
ABI in include/spray_gpu.h); this package only binds it.
"""
from . import _lib
from .api import (Context, SprayError, default_config, scenegen_domain, scenegen_write, write_ply_info,
                  write_ppm)

__all__ = ["Context", "SprayError", "default_config", "scenegen_domain", "scenegen_write", "write_ply_info", "write_ppm",
           "_lib"]
