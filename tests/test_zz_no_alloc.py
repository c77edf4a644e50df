"""psy_trace_noalloc against psy_trace_async through the C ABI: the entry
point that allocates nothing writes the same hit records as the tracer's
stream-ordered trace, on two calls in a row (the second with fewer rays, so
the shared scratch counter must be reset between calls), on the default
stream and on a caller's stream."""

import ctypes as C

import numpy as np
import pytest

from psychopath_amd import _lib, scenes
from psychopath_amd.tracer import Assembly, DeviceHitRecords, DeviceRays, Mesh, RayBatch, Tracer

pytestmark = pytest.mark.gpu


def _rays(lo, hi, n, seeds):
    r = scenes.incoherent_rays(lo, hi, n, seeds=seeds)
    return RayBatch(r["ox"], r["oy"], r["oz"], r["dx"], r["dy"], r["dz"], r["time"], r["max_t"])


def _noalloc(scene, rays, hits, tuning=None, stream=None):
    rs, hs = rays.soa(), hits.soa()
    t = C.byref(tuning) if tuning is not None else None
    _lib.check("psy_trace_noalloc",
               _lib.lib().psy_trace_noalloc(scene.handle, C.byref(rs), rays.n, C.byref(hs), t, stream))


def _hip():
    """The HIP runtime, for a caller-owned stream."""
    hip = C.CDLL("libamdhip64.so")
    for name in ("hipStreamCreate", "hipStreamSynchronize", "hipStreamDestroy", "hipDeviceSynchronize"):
        getattr(hip, name).restype = C.c_int
    return hip


def test_noalloc_matches_tracer_on_two_calls():
    v, f = scenes.displaced_sphere(7)
    asm = Assembly()
    asm.add_mesh(Mesh(v, f))
    asm.add_instance(0)
    scene = asm.build()
    tracer = Tracer(scene, max_batch=1 << 16)  # uploads the scene
    lo, hi = v.min(0), v.max(0)
    big = DeviceRays(_rays(lo, hi, 150_000, (2, 3, 4)))
    small = DeviceRays(_rays(lo, hi, 40_000, (5, 6, 7)))

    want_big, want_small = DeviceHitRecords(big.n), DeviceHitRecords(small.n)
    tracer.trace_device(big, want_big, sync=False)
    tracer.sync()
    tracer.trace_device(small, want_small, sync=False)
    tracer.sync()
    wb, ws = want_big.records(), want_small.records()
    assert (wb["kind"] == _lib.PSY_HIT_SURFACE).sum() > 10_000

    # Call 1: the big batch on the default stream with the default shape.
    got_big = DeviceHitRecords(big.n)
    _noalloc(scene, big, got_big)
    hip = _hip()
    assert hip.hipDeviceSynchronize() == 0
    assert np.array_equal(got_big.records(), wb)

    # Call 2: the smaller batch on a caller's stream, cooperative fetch shape.
    got_small = DeviceHitRecords(small.n)
    tt = _lib.TraceTuning(_lib.PSY_KERNEL_PERSISTENT_COOP, 8, 16, 16, 4, 16)
    stream = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(stream)) == 0
    try:
        _noalloc(scene, small, got_small, tt, stream)
        assert hip.hipStreamSynchronize(stream) == 0
    finally:
        hip.hipStreamDestroy(stream)
    assert np.array_equal(got_small.records(), ws)

    # Errors: more rays than the counter holds, and the simple kernel.
    rs, hs = small.soa(), got_small.soa()
    L = _lib.lib()
    assert L.psy_trace_noalloc(scene.handle, C.byref(rs), _lib.PSY_NOALLOC_MAX_RAYS + 1, C.byref(hs), None,
                               None) == _lib.PSY_E_RANGE
    simple = _lib.TraceTuning(_lib.PSY_KERNEL_SIMPLE, 1, 1, 8, 4, 0)
    assert L.psy_trace_noalloc(scene.handle, C.byref(rs), small.n, C.byref(hs), C.byref(simple),
                               None) == _lib.PSY_E_INVALID
