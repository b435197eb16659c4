"""MI355X (gfx950) 3D smoke step: the ZQ_SmokeSimulation3D hot path on hand-written HIP kernels.

The compute path is libzq_smoke3d.so (C ABI: include/zq_smoke3d.h).  This package
is its Python front end:
  stages   reference-compatible host-pointer stage calls (numpy in, numpy out)
  sim      device-resident simulation handle
"""
from . import _lib

_lib.lib()  # load libzq_smoke3d.so now: no library, no package (there is no fallback path)
from ._lib import ZQ3DError, last_error  # noqa: F401
from .sim import SmokeSim3D, default_config, noise_field  # noqa: F401
from . import stages  # noqa: F401

__all__ = ["SmokeSim3D", "default_config", "noise_field", "stages", "ZQ3DError", "last_error"]
