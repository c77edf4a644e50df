"""Pins the oracle's GGX closure, sphere- and distant-light samplers and the
light tree's selection with checks the mathematics fixes, chosen so that a
plausible slip (a dropped pi, a squared roughness taken once, a wrong lerp
direction, a mis-scaled cone, an unnormalised or mis-rescaled selection)
fails one of them.  None of them retypes the oracle's formulas: each compares
against a textbook statement evaluated independently in float64 (the GGX
normal distribution's normalisation and Smith masking term, Schlick's
approximation, solid angles of cones and spheres, the irradiance of a sphere
of uniform radiance), or against a property the algorithm must have
(sampled densities that equal the stated pdfs, selection frequencies that
equal the stated probabilities).

Reference code: src/shading/surface_closure.rs:411-633 (GGX),
src/light/sphere_light.rs:51-192, src/light/distant_disk_light.rs:59-85,
src/accel/light_tree.rs:135-201 (select) and :211-372 (build).
"""
import numpy as np
import pytest
from scipy import stats

from oracle import psyref

GGX, LAMBERT = 1, 0
N = np.array([0.0, 0.0, 1.0])


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _ggx_d_textbook(cos_h, alpha):
    """Trowbridge-Reitz / GGX normal distribution (Walter et al. 2007, eq. 33)."""
    a2 = alpha * alpha
    t = (a2 - 1.0) * cos_h * cos_h + 1.0
    return a2 / (np.pi * t * t)


def _smith_g1_textbook(cos_v, alpha):
    """Smith masking for GGX (Walter et al. 2007, eq. 34)."""
    tan2 = (1.0 - cos_v * cos_v) / (cos_v * cos_v)
    return 2.0 / (1.0 + np.sqrt(1.0 + alpha * alpha * tan2))


def _retro_d(rough, theta):
    """D(cos theta) read off the oracle's evaluate(): with the outgoing
    direction equal to the incoming one reversed, the half vector is that
    direction, and the reference's pdf is D(n.h) / (pi n.i)."""
    aa = np.stack([np.sin(theta), np.zeros_like(theta), np.cos(theta)], 1)
    _, pdf = psyref.closure_evaluate(GGX, -aa, aa, N, N, roughness=rough)
    return pdf.astype(np.float64) * np.cos(theta) * np.pi


# ---- GGX -------------------------------------------------------------------

@pytest.mark.parametrize("rough", [0.05, 0.2, 0.5, 0.8, 1.0])
def test_ggx_distribution_is_normalised(rough):
    """Integral over the hemisphere of D(h) (n.h) d(omega_h) = 1."""
    th = np.linspace(0.0, np.pi / 2, 40001)[:-1]
    d = _retro_d(rough, th)
    integral = np.trapezoid(d * np.cos(th) * np.sin(th) * 2 * np.pi, th)
    assert abs(integral - 1.0) < 2e-5, integral


@pytest.mark.parametrize("rough", [0.05, 0.3, 0.7, 1.0])
def test_ggx_distribution_matches_textbook(rough):
    th = np.linspace(0.0, 1.45, 300)
    d = _retro_d(rough, th)
    want = _ggx_d_textbook(np.cos(th), rough)
    # float32 directions: the steep low-roughness lobe amplifies their last bit
    np.testing.assert_allclose(d, want, rtol=1e-4, atol=1e-30)


@pytest.mark.parametrize("rough,beta", [(0.1, 0.0), (0.3, 0.6), (0.6, 1.1), (0.9, 0.3)])
def test_ggx_sampled_half_vectors_follow_d_cos(rough, beta):
    """sample() draws half vectors with density D(h) (n.h), the density whose
    normalisation the previous test checks: the CDF of the sampled polar
    angle, built by quadrature of the oracle's own D, is uniform on the
    samples (KS test), and the azimuth is uniform."""
    rng = np.random.default_rng(7)
    uv = rng.random((20000, 2)).astype(np.float32)
    aa = np.array([np.sin(beta), 0.0, np.cos(beta)])
    out, _, _ = psyref.closure_sample(GGX, -aa, N, N, uv, roughness=rough)
    h = _unit(aa[None, :] + _unit(out))
    # Reflection about h and about -h is the same: a half vector facing away
    # from the incoming direction comes back flipped.
    h *= np.sign(h[:, 2])[:, None]
    th_s = np.arccos(np.clip(h[:, 2], -1, 1))
    grid = np.linspace(0.0, np.pi / 2, 20001)
    dens = _retro_d(rough, grid[:-1]) * np.cos(grid[:-1]) * np.sin(grid[:-1]) * 2 * np.pi
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(grid[:-1]))])
    cdf = np.concatenate([cdf, [1.0]])
    u_back = np.interp(th_s, grid, cdf)
    assert stats.kstest(u_back, "uniform").pvalue > 1e-3
    # Azimuth of h about the normal (the frame of zup_to_vec is a rotation
    # about n, so uniformity is frame independent).
    phi = np.arctan2(h[:, 1], h[:, 0])
    assert stats.kstest((phi + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3


def test_ggx_sample_reflects_about_the_half_vector():
    rng = np.random.default_rng(11)
    uv = rng.random((2000, 2)).astype(np.float32)
    inc = _unit([0.3, -0.2, -1.0])
    out, filt, pdf = psyref.closure_sample(GGX, inc, N, N, uv, roughness=0.4)
    # |out| = |inc|, and out - inc is parallel to the sampled half vector,
    # which lies in the upper hemisphere (up to the sign that reflection
    # does not see), so out is inc mirrored about h.
    np.testing.assert_allclose(np.linalg.norm(out, axis=1), 1.0, rtol=1e-5)
    h = _unit(_unit(out) - inc[None, :])
    h *= np.sign(h[:, 2])[:, None]
    mirrored = inc[None, :] - 2 * h * (h @ inc)[:, None]
    np.testing.assert_allclose(out, mirrored, atol=3e-6)
    # The returned pdf is evaluate()'s for the returned direction.
    f2, p2 = psyref.closure_evaluate(GGX, np.broadcast_to(inc, out.shape), out, N, N, roughness=0.4)
    np.testing.assert_array_equal(pdf, p2)
    np.testing.assert_array_equal(filt, f2)


@pytest.mark.parametrize("beta", [0.2, 0.7, 1.2])
@pytest.mark.parametrize("fresnel", [0.0, 0.5, 1.0])
def test_ggx_mirror_weight_is_schlick_times_smith(beta, fresnel):
    """At the mirror configuration (h = n) the weight filter/pdf of a
    reflection is F(h.o) G1(i) G1(o), with F Schlick's approximation
    f + (1-f)(1 - cos)^5 faded toward the plain colour by 1 - fresnel and
    G1 the Smith term."""
    rough = 0.35
    aa = np.array([np.sin(beta), 0.0, np.cos(beta)])
    bb = np.array([-np.sin(beta), 0.0, np.cos(beta)])
    xyz = (0.35, 0.25, 0.1)
    spec = np.zeros(4, np.float32)
    psyref.lib().psyref_xyz_to_spectrum(*xyz, 550.0, psyref.fptr(spec))
    filt, pdf = psyref.closure_evaluate(GGX, -aa, bb, N, N, color=(0,) + xyz, roughness=rough, fresnel=fresnel)
    w = filt[0].astype(np.float64) / float(pdf[0])
    c = np.cos(beta)
    f = spec.astype(np.float64)
    schlick = f + (1 - f) * (1 - c) ** 5
    col = fresnel * schlick + (1 - fresnel) * f
    want = col * _smith_g1_textbook(c, rough) ** 2
    np.testing.assert_allclose(w, want, rtol=3e-5)


def test_ggx_furnace_tends_to_four_over_pi():
    """White furnace of the reference's GGX: the reflected energy
    integral of filter over the outgoing hemisphere.  The reference's filter
    is F D G1 G2 / (pi n.i) (surface_closure.rs:513-518), where a textbook
    microfacet BRDF times n.o would be F D G / (4 n.i); its furnace therefore
    tends to 4/pi, not 1, as the surface turns smooth, at every incidence, and
    falls as roughness (and with it masking) grows.  The integral is taken
    over half vectors drawn by the textbook GGX inverse CDF with the reflection
    Jacobian 4 (o.h)."""
    def furnace(rough, beta, n=400):
        aa = np.array([np.sin(beta), 0.0, np.cos(beta)])
        u = (np.arange(n) + 0.5) / n
        U, V = (x.ravel() for x in np.meshgrid(u, u))
        a2 = rough * rough
        c = np.sqrt((1 - U) / (1 + (a2 - 1) * U))
        s = np.sqrt(1 - c * c)
        h = np.stack([np.cos(2 * np.pi * V) * s, np.sin(2 * np.pi * V) * s, c], 1)
        out = -aa[None, :] + 2 * h * (h @ aa)[:, None]
        f, _ = psyref.closure_evaluate(GGX, -aa, out, N, N, roughness=rough)
        w = 4 * np.abs((out * h).sum(1)) / (_ggx_d_textbook(c, rough) * c)
        return float(np.mean(f[:, 0] * w))

    for beta in (0.0, 0.5, 1.0):
        assert abs(furnace(0.02, beta) - 4 / np.pi) < 0.005 * 4 / np.pi, beta
    vals = [furnace(r, 0.0) for r in (0.05, 0.2, 0.5, 0.8)]
    assert all(a > b for a, b in zip(vals, vals[1:])), vals


def test_ggx_sampled_bounce_never_gains_energy():
    """A sampled bounce carries filter/pdf = F G1 G2: at most the colour's
    spectral value (E-white upsamples to 1 within the Meng tables' error)."""
    white = np.zeros(4, np.float32)
    psyref.lib().psyref_xyz_to_spectrum(1.0, 1.0, 1.0, 550.0, psyref.fptr(white))
    assert abs(white[0] - 1) < 1e-3
    rng = np.random.default_rng(5)
    uv = rng.random((5000, 2)).astype(np.float32)
    for rough in (0.05, 0.4, 0.9):
        for beta in (0.1, 0.8, 1.3):
            aa = np.array([np.sin(beta), 0.0, np.cos(beta)])
            _, filt, pdf = psyref.closure_sample(GGX, -aa, N, N, uv, roughness=rough)
            ok = pdf > 0
            w = filt[ok, 0] / pdf[ok]
            assert np.all(w <= white[0] * (1 + 1e-6)) and np.all(w >= 0), (rough, beta, w.max())


def test_ggx_below_horizon_and_mirror_cases():
    aa = _unit([0.3, 0.0, 1.0])
    # Outgoing direction below the surface: nothing.
    f, p = psyref.closure_evaluate(GGX, -aa, [0.2, 0.1, -1.0], N, N, roughness=0.3)
    assert p[0] == 0 and np.all(f == 0)
    # Roughness 0 is a perfect mirror: colour times Fresnel, pdf 0.
    f, p = psyref.closure_evaluate(GGX, -aa, aa * [-1, -1, 1], N, N, roughness=0.0)
    assert p[0] == 0 and f[0, 0] > 0


# ---- sphere light ----------------------------------------------------------

def _translate(c):
    return np.array([1, 0, 0, -c[0], 0, 1, 0, -c[1], 0, 0, 1, -c[2]], np.float32)


def _strat(n, seed=0):
    rng = np.random.default_rng(seed)
    k = int(np.sqrt(n))
    g = (np.arange(k) + 0.5) / k
    U, V = np.meshgrid(g, g)
    uv = np.stack([U.ravel(), V.ravel()], 1)
    return np.clip(uv + rng.uniform(-0.5 / k, 0.5 / k, uv.shape), 0, 0.9999999).astype(np.float32)


# The reference builds the sampling frame with the Duff et al. orthonormal
# basis of the UNnormalised vector from the shading point to the centre
# (sphere_light.rs:111, math.rs:26-34); that basis is orthonormal only for a
# unit vector or one along +z.  Where it is, every property of uniform cone
# sampling holds exactly and is checked below; elsewhere the oracle keeps the
# reference's skewed frame (the last test shows it does).
@pytest.mark.parametrize("r,d,axis", [(0.3, 2.0, "z"), (0.5, 0.8, "z"), (0.05, 7.0, "z"), (1.0, 1.0001, "z"),
                                      (0.3, 1.0, "off"), (0.9, 1.0, "off"), (0.02, 1.0, "off")])
def test_sphere_light_samples_uniform_cone(r, d, axis):
    c = np.array([0.4, -1.2, 2.0])
    dirn = np.array([0.0, 0.0, -1.0]) if axis == "z" else _unit([0.3, 0.5, -0.8])
    arr = c + d * dirn
    uv = _strat(40000)
    s = psyref.sphere_light_sample(r, arr, uv, xform=_translate(c))
    pt = s["point"].astype(np.float64)
    # On the sphere.
    np.testing.assert_allclose(np.linalg.norm(pt - c, axis=1), r, rtol=2e-5)
    # Inside the cap seen from the shading point: cos(angle at the centre
    # between the point and the viewer) >= r/d.
    cosg = ((pt - c) * _unit(arr - c)).sum(1) / r
    assert np.all(cosg >= r / d - 2e-4)
    # The direction to each point is the u-th cone direction: its cosine to
    # the axis is 1 - u (1 - cos theta_max).
    ctm = np.sqrt(max(1 - (r / d) ** 2, 0.0))
    cos_t = (_unit(pt - arr) * _unit(c - arr)).sum(1)
    # (float32 points: the cosine's error grows as the viewer nears the surface)
    np.testing.assert_allclose(cos_t, 1 - uv[:, 0] * (1 - ctm), atol=3e-5 + 6e-8 * r / (d - r))
    # Uniform azimuth about the axis.
    ax = _unit(c - arr)
    e1 = _unit(np.cross(ax, [0.3, 1.0, 0.2]))
    e2 = np.cross(ax, e1)
    w = _unit(pt - arr)
    phi = np.arctan2(w @ e2, w @ e1)
    assert stats.kstest((phi + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3
    # pdf of a uniform cone of half-angle theta_max, and sample_pdf agrees.
    pdf = 1 / (2 * np.pi * (1 - ctm))
    np.testing.assert_allclose(s["pdf"], pdf, rtol=1e-5)
    np.testing.assert_allclose(psyref.sphere_light_pdf(r, arr, xform=_translate(c)), pdf, rtol=1e-5)
    # Outward normals.
    np.testing.assert_allclose(s["normal"], _unit(pt - c), atol=1e-4)


def test_sphere_light_keeps_the_reference_frame_off_axis():
    """Off the +z axis and away from unit distance the reference's frame is
    not orthonormal: points stay on the sphere with the cone's pdf, but some
    leave the cap visible from the shading point.  An oracle that normalised
    the vector first would put them all inside."""
    r, d = 0.3, 2.0
    c = np.array([0.4, -1.2, 2.0])
    arr = c + d * _unit([0.3, 0.5, -0.8])
    s = psyref.sphere_light_sample(r, arr, _strat(10000), xform=_translate(c))
    pt = s["point"].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(pt - c, axis=1), r, rtol=2e-5)
    np.testing.assert_allclose(s["pdf"], 1 / (2 * np.pi * (1 - np.sqrt(1 - (r / d) ** 2))), rtol=1e-5)
    cosg = ((pt - c) * _unit(arr - c)).sum(1) / r
    assert np.mean(cosg < r / d - 1e-3) > 0.05


def test_sphere_light_irradiance_closed_form():
    """A Lambertian point facing a sphere of uniform radiance L has
    irradiance E = pi L sin^2(theta_max) = pi L (r/d)^2; the estimator
    mean(L cos / pdf) over the light's samples must reproduce it."""
    r, d = 0.25, 1.7
    c = np.zeros(3)
    arr = np.array([0.0, 0.0, -d])
    uv = _strat(90000, 3)
    s = psyref.sphere_light_sample(r, arr, uv, color=(0, 1.0, 1.0, 1.0))
    L = s["color"][:, 0].astype(np.float64)
    cos = _unit(s["point"].astype(np.float64) - arr)[:, 2]
    est = np.mean(L * cos / s["pdf"])
    want = np.pi * L[0] * (r / d) ** 2
    assert abs(est - want) < 1e-4 * want, (est, want)
    # Radiance is the colour over the sphere's area.
    spec = np.zeros(4, np.float32)
    psyref.lib().psyref_xyz_to_spectrum(1.0, 1.0, 1.0, 550.0, psyref.fptr(spec))
    np.testing.assert_allclose(s["color"][0], spec / (4 * np.pi * r * r), rtol=1e-6)


def test_sphere_light_from_inside_samples_whole_sphere():
    r = 2.0
    c = np.array([1.0, 0.0, 0.0])
    arr = c + np.array([0.3, -0.2, 0.5])
    uv = _strat(20000, 4)
    s = psyref.sphere_light_sample(r, arr, uv, xform=_translate(c))
    np.testing.assert_allclose(np.linalg.norm(s["point"] - c, axis=1), r, rtol=2e-5)
    np.testing.assert_allclose(s["pdf"], 1 / (4 * np.pi), rtol=1e-6)
    np.testing.assert_allclose(psyref.sphere_light_pdf(r, [arr], xform=_translate(c)), 1 / (4 * np.pi), rtol=1e-6)


def test_sphere_light_under_rotation_and_scale():
    """The light's transform moves and scales the sphere: with a rotation
    and a uniform scale s (world-to-object), samples lie on a sphere of
    radius r/s around the transformed centre."""
    ang = 0.7
    R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    sc = 2.0
    c = np.array([0.5, 1.0, -3.0])
    M = sc * R
    t = -M @ c
    xf = np.array([*M[0], t[0], *M[1], t[1], *M[2], t[2]], np.float32)
    r = 0.4
    arr = c + np.array([0.0, 2.5, 0.7])
    uv = _strat(10000, 5)
    s = psyref.sphere_light_sample(r, arr, uv, xform=xf)
    np.testing.assert_allclose(np.linalg.norm(s["point"] - c, axis=1), r / sc, rtol=5e-5)


# ---- distant disk light ----------------------------------------------------

@pytest.mark.parametrize("radius", [0.01, 0.1, 0.5, 1.2])
def test_distant_disk_light_samples_cone(radius):
    direction = _unit([0.2, -0.4, -1.0])
    uv = _strat(40000, 6)
    col, sv, pdf = psyref.distant_light_sample(radius, direction, uv)
    cos_t = (_unit(sv) * -direction).sum(1)
    ctm = np.cos(radius)
    # Shadow rays point back toward the light, inside its cone, at the
    # cone's u-th cosine.
    assert np.all(cos_t >= ctm - 1e-5)
    np.testing.assert_allclose(cos_t, 1 - uv[:, 0] * (1 - ctm), atol=2e-6 / max(1 - ctm, 1e-4) ** 0.5)
    solid = 2 * np.pi * (1 - ctm)
    np.testing.assert_allclose(pdf, 1 / solid, rtol=1e-5)
    spec = np.zeros(4, np.float32)
    psyref.lib().psyref_xyz_to_spectrum(1.0, 1.0, 1.0, 550.0, psyref.fptr(spec))
    np.testing.assert_allclose(col[0], spec / solid, rtol=1e-5)
    # Uniform over the cone's solid angle: the azimuth about the axis too.
    ax = -direction
    e1 = _unit(np.cross(ax, [1.0, 0, 0]))
    e2 = np.cross(ax, e1)
    phi = np.arctan2(_unit(sv) @ e2, _unit(sv) @ e1)
    assert stats.kstest((phi + np.pi) / (2 * np.pi), "uniform").pvalue > 1e-3


# ---- light tree selection --------------------------------------------------

def _sphere_boxes(centres, radii):
    centres = np.asarray(centres, np.float64)
    radii = np.asarray(radii, np.float64)[:, None]
    return np.concatenate([centres - radii, centres + radii], 1).astype(np.float32)


def test_light_tree_small_sets_select_by_irradiance():
    """With at most four lights the collapsed tree holds every light as a
    child of the root, so a light's probability is its weight over the sum.
    For a Lambertian point with every light fully above its horizon, the
    weight energy / r^2 * estimate reduces to energy * cos / dist^2 (the
    sphere estimate is cos * sin^2(theta_max) and sin^2(theta_max) =
    r^2/dist^2 for the box's bounding sphere), whatever the radii."""
    rng = np.random.default_rng(21)
    pos = np.array([0.1, -0.2, 0.0])
    for trial in range(12):
        n = int(rng.integers(2, 5))
        centres = pos + np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(2, 3, n)], 1)
        radii = rng.uniform(0.1, 0.3, n)
        energy = rng.uniform(0.5, 10, n)
        boxes = _sphere_boxes(centres, radii)
        us = ((np.arange(4000) + 0.5) / 4000).astype(np.float32)
        idx, prob, _ = psyref.light_tree_select(boxes, energy, us, inc=[0, 0, -1.0], pos=pos, nor=[0, 0, 1.0])
        b = boxes.astype(np.float64)
        d = 0.5 * (b[:, :3] + b[:, 3:]) - pos
        dist2 = (d * d).sum(1)
        w = energy * (d[:, 2] / np.sqrt(dist2)) / dist2
        want = w / w.sum()
        for i in range(n):
            got = prob[idx == i]
            assert got.size > 0
            # float32: sin^2(theta_max) is formed as 1 - cos^2, which
            # cancels digits for small lights
            np.testing.assert_allclose(got, want[i], rtol=1e-4, err_msg=f"trial {trial} light {i}")


@pytest.mark.parametrize("kind,rough,seed", [(LAMBERT, 0.0, 1), (GGX, 0.3, 2), (LAMBERT, 0.0, 3)])
def test_light_tree_selection_frequencies_match_probabilities(kind, rough, seed):
    """On a deep tree (60 lights), stratified numbers n in [0, 1) select each
    light with the frequency of its stated probability; a light's stated
    probability does not depend on n; the probabilities of all lights sum to
    one; and the whittled number, rescaled at each level, runs from 0 to 1
    linearly over each light's interval with slope 1/prob."""
    rng = np.random.default_rng(seed)
    n = 60
    centres = rng.uniform(-5, 5, (n, 3))
    radii = rng.uniform(0.05, 0.5, n)
    energy = rng.uniform(0.1, 20, n)
    boxes = _sphere_boxes(centres, radii)
    m = 400000
    us = ((np.arange(m) + 0.5) / m).astype(np.float32)
    pos = np.array([0.0, 0.0, -6.0])
    nor = _unit([0.1, 0.2, 1.0])
    idx, prob, wh = psyref.light_tree_select(boxes, energy, us, inc=_unit([0.3, 0, 1.0]) * -1, pos=pos,
                                             nor=nor, kind=kind, roughness=rough)
    seen = np.unique(idx)
    total = 0.0
    for i in seen:
        sel = idx == i
        p = prob[sel]
        assert np.all(p == p[0]), i
        total += float(p[0])
        # Frequency equals probability (stratified: within a few strata).
        assert abs(sel.sum() / m - p[0]) <= 3.0 / m + 1e-5 * p[0], (i, sel.sum() / m, p[0])
        # The selected n form one interval; the whittled number maps it to
        # [0, 1) with slope 1/prob.
        k = np.nonzero(sel)[0]
        assert k[-1] - k[0] + 1 == k.size, "light chosen on a non-contiguous set of n"
        if k.size > 20:
            slope = np.polyfit(us[k].astype(np.float64), wh[k].astype(np.float64), 1)[0]
            assert abs(slope * p[0] - 1) < 1e-3, (i, slope, 1 / p[0])
            assert wh[k].min() >= -1e-6 and wh[k].max() <= 1 + 1e-5
    assert abs(total - 1.0) < 1e-5, total
    # Lights whose probability is large enough to be hit must be hit.
    assert seen.size >= 10


def test_light_tree_uniform_when_every_estimate_is_zero():
    """Lights all below a Lambertian point's horizon give every child weight
    zero, and select() falls back to equal probabilities at each level, so
    each light's probability is the product of 1/arity along its path."""
    rng = np.random.default_rng(4)
    n = 20
    centres = np.stack([rng.uniform(-3, 3, n), rng.uniform(-3, 3, n), rng.uniform(-6, -3, n)], 1)
    boxes = _sphere_boxes(centres, rng.uniform(0.05, 0.2, n))
    energy = rng.uniform(1, 5, n)
    m = 200000
    us = ((np.arange(m) + 0.5) / m).astype(np.float32)
    idx, prob, _ = psyref.light_tree_select(boxes, energy, us, inc=[0, 0, -1.0], pos=[0, 0, 0.0], nor=[0, 0, 1.0])
    # The tree's structure, from its dump: probability = product of 1/arity.
    dump = psyref.light_tree_dump(boxes, energy)
    want = {}

    def walk(k, p):
        leaf, val = dump[k, 0], dump[k, 1]
        if leaf:
            want[int(val)] = p
            return k + 1
        nxt = k + 1
        for _ in range(val):
            nxt = walk(nxt, p / val)
        return nxt

    assert walk(0, 1.0) == len(dump)
    assert sorted(want) == list(range(n))
    for i in range(n):
        np.testing.assert_allclose(prob[idx == i], want[i], rtol=1e-6)


def test_light_tree_structure():
    """Each light is one leaf; arity 2..8 after the collapse (light_tree.rs:
    296-372); a node's energy is the sum of its children's; a child's box lies
    inside its parent's."""
    rng = np.random.default_rng(9)
    for n in (1, 2, 3, 7, 8, 9, 50, 300):
        centres = rng.uniform(-10, 10, (n, 3))
        boxes = _sphere_boxes(centres, rng.uniform(0.01, 1.0, n))
        energy = rng.uniform(0.1, 3, n).astype(np.float32)
        dump = psyref.light_tree_dump(boxes, energy)
        e = dump[:, 2].view(np.float32)
        bx = dump[:, 3:].view(np.float32)
        leaves = []

        def walk(k):
            leaf, val = dump[k, 0], dump[k, 1]
            if leaf:
                leaves.append(int(val))
                assert e[k] == energy[val]
                np.testing.assert_array_equal(bx[k], boxes[val])
                return k + 1
            assert 2 <= val <= 8, val
            nxt = k + 1
            s = 0.0
            for _ in range(val):
                assert np.all(bx[nxt, :3] >= bx[k, :3]) and np.all(bx[nxt, 3:] <= bx[k, 3:])
                s += float(e[nxt])
                nxt = walk(nxt)
            assert abs(e[k] - s) <= 1e-5 * s
            return nxt

        assert walk(0) == len(dump)
        assert sorted(leaves) == list(range(n))
