"""MI355X-native engine for sphere's DEM particle timestep.

`sim` mirrors the reference's Python front end (python/sphere.py); the
timestep runs in hand-written HIP kernels behind the C ABI of
include/sphere_hip.h (libsphere_hip.so).
"""

from .sim import V_sphere, cleanup, run, sim, status  # noqa: F401
from .binfmt import VERSION  # noqa: F401

__all__ = ["sim", "run", "status", "cleanup", "V_sphere", "VERSION"]
