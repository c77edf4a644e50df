"""psychopath_amd — MI355X (gfx950) ray-tracing path for the psychopath renderer.

The compute path is libpsy_rt.so (HIP kernels behind the C ABIs of
include/psy_rt.h, psy_render.h, psy_file.h and psy_image.h); this package is
the host-side mirror of the reference's interfaces over them:

    tracer    Assembly / Tracer / RayBatch: ray batches in, hit records out
    render    Renderer (wavefront and one-launch forms), Server (continuous
              batching of render requests, streamed rows)
    psyfile   .psy scene files in and out
    film      PPM / PNG / EXR writers and the Blender bucket protocol
    scenes    the seeded benchmark workloads
    __main__  the reference's command line (python -m psychopath_amd -i scene.psy)
"""

from ._lib import PsyError, lib  # noqa: F401
from .tracer import Assembly, DeviceHits, DeviceRays, Hits, Mesh, RayBatch, Scene, Tracer  # noqa: F401

__all__ = ["Assembly", "Mesh", "Scene", "Tracer", "RayBatch", "Hits", "DeviceRays", "DeviceHits", "PsyError", "lib"]
