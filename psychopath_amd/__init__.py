"""psychopath_amd — MI355X (gfx950) ray-tracing path for the psychopath renderer.

The compute path is libpsy_rt.so (HIP kernels behind the C ABI of
include/psy_rt.h); this package is the host-side mirror of the reference's
tracing interface over that ABI.
"""

from ._lib import PsyError, lib  # noqa: F401
from .tracer import Assembly, DeviceHits, DeviceRays, Hits, Mesh, RayBatch, Scene, Tracer  # noqa: F401

__all__ = ["Assembly", "Mesh", "Scene", "Tracer", "RayBatch", "Hits", "DeviceRays", "DeviceHits", "PsyError", "lib"]
