"""Independent brute-force references in float64 (test helpers).

Möller-Trumbore ray-triangle intersection and slab clipping, evaluated in
double precision over every primitive: these share no code or operation
order with the oracle's Woop test or BVH traversal.
"""

import numpy as np


def mt_intersect(o, d, v0, v1, v2, eps=0.0):
    """Rays o,d [R,3] against triangles v0,v1,v2 [T,3] (float64).
    Returns t [R,T] (inf where missed) and barycentrics (w, u, v) [R,T,3]:
    w for v0, u for v1, v for v2."""
    o = o[:, None, :].astype(np.float64)
    d = d[:, None, :].astype(np.float64)
    e1 = (v1 - v0)[None].astype(np.float64)
    e2 = (v2 - v0)[None].astype(np.float64)
    p = np.cross(d, e2)
    det = np.einsum("rtk,rtk->rt", e1, p)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / det
        s = o - v0[None]
        u = np.einsum("rtk,rtk->rt", s, p) * inv
        q = np.cross(s, e1)
        v = np.einsum("rtk,rtk->rt", d, q) * inv
        t = np.einsum("rtk,rtk->rt", e2, q) * inv
    ok = (det != 0) & (u >= -eps) & (v >= -eps) & (u + v <= 1 + eps) & (t > 0)
    t = np.where(ok, t, np.inf)
    bary = np.stack([1 - u - v, u, v], axis=-1)
    return t, bary


def closest_hits(o, d, max_t, verts, tris, chunk=256):
    """Brute-force closest hit per ray over every triangle.
    Returns (t_best, prim_best, t_second) with t = inf for misses."""
    v0, v1, v2 = (verts[tris[:, k]].astype(np.float64) for k in range(3))
    R = len(o)
    tb = np.full(R, np.inf)
    t2 = np.full(R, np.inf)
    pb = np.full(R, -1, np.int64)
    for a in range(0, R, chunk):
        t, _ = mt_intersect(o[a:a + chunk], d[a:a + chunk], v0, v1, v2)
        t = np.where(t < max_t[a:a + chunk, None], t, np.inf)
        idx = np.argsort(t, axis=1)[:, :2]
        rows = np.arange(len(t))
        tb[a:a + chunk] = t[rows, idx[:, 0]]
        pb[a:a + chunk] = np.where(np.isfinite(tb[a:a + chunk]), idx[:, 0], -1)
        if t.shape[1] > 1:
            t2[a:a + chunk] = t[rows, idx[:, 1]]
    return tb, pb, t2


def slab_f64(lo, hi, o, d, max_t):
    """Exact-arithmetic-ish slab test in float64: [R,3] rays vs one box each."""
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d.astype(np.float64)
        t1 = (lo - o) * inv
        t2 = (hi - o) * inv
    tn = np.nanmax(np.maximum(np.minimum(t1, t2), -np.inf), axis=1)
    tf = np.nanmin(np.maximum(t1, t2), axis=1)
    tn = np.maximum(tn, 0.0)
    tf = np.minimum(tf, max_t)
    return tn, tf
