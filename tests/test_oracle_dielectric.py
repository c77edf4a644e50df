"""The rough dielectric closure (PSY_SHADER_DIELECTRIC), which the reference
does not have (its closures are Lambert, GGX and Emit,
surface_closure.rs:19-30): the oracle's restatement is pinned by what the
physics fixes, so that the GPU renders compared with it are checked against
the specification rather than against a shared misreading.

* Fresnel: at the exact mirror configuration the reflection density is
  F D(1) / (4 cos i) with D(1) = 1 / (pi alpha^2), so F read back from it
  must equal the textbook unpolarised Fresnel reflectance, including total
  internal reflection from inside beyond the critical angle.
* Snell's law: refracted samples of a nearly smooth surface satisfy
  eta_i sin(i) = eta_o sin(t) in the plane of incidence.
* Densities: the evaluate density integrated over the sphere (uniform
  Monte Carlo) equals the probability that sample() returns a direction,
  and the reflected share equals its integral over the reflection side;
  a wrong Jacobian (eta_o^2 |o.m| / (eta_i i.m + eta_o o.m)^2) fails this.
* Energy: a white dielectric neither creates energy (mean weight <= 1) nor
  loses much when smooth, entering or leaving.
* Reflection is reciprocal.
"""

import math

import numpy as np
import pytest

from oracle import psyref

K = psyref.SHADER_DIELECTRIC
WHITE = (0, 1.0, 1.0, 1.0)
N = np.array([0.0, 0.0, 1.0], np.float32)


def fresnel_textbook(cos_i, eta_i, eta_t):
    s2t = (eta_i / eta_t) ** 2 * (1.0 - cos_i * cos_i)
    if s2t >= 1.0:
        return 1.0
    cos_t = math.sqrt(1.0 - s2t)
    rs = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t)
    rp = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t)
    return 0.5 * (rs * rs + rp * rp)


def incoming(theta, inside=False):
    """A ray direction arriving at angle theta from the +z side (or from
    below, travelling up, when inside)."""
    d = np.array([math.sin(theta), 0.0, -math.cos(theta)], np.float32)
    if inside:
        d[2] = -d[2]
    return d


@pytest.mark.parametrize("inside", [False, True])
@pytest.mark.parametrize("theta_deg", [0.0, 20.0, 40.0, 45.0, 60.0, 80.0])
def test_fresnel_at_the_mirror_direction(theta_deg, inside):
    ior, alpha = 1.5, 0.1
    th = math.radians(theta_deg)
    inc = incoming(th, inside)
    out = inc.copy()
    out[2] = -out[2]                      # mirror about the z axis
    _, pdf = psyref.closure_evaluate(K, inc, out, N, N, color=WHITE, roughness=alpha, ior=ior)
    d1 = 1.0 / (math.pi * alpha * alpha)
    f = float(pdf[0]) * 4.0 * math.cos(th) / d1
    eta_i, eta_t = (1.5, 1.0) if inside else (1.0, 1.5)
    want = fresnel_textbook(math.cos(th), eta_i, eta_t)
    assert f == pytest.approx(want, rel=2e-3, abs=2e-5), (theta_deg, inside)


@pytest.mark.parametrize("inside", [False, True])
def test_snell_law_for_refracted_samples(inside):
    ior = 1.5
    rng = np.random.default_rng(5)
    eta_i, eta_t = (ior, 1.0) if inside else (1.0, ior)
    for theta_deg in (10.0, 30.0, 40.0, 60.0):
        th = math.radians(theta_deg)
        if math.sin(th) * eta_i / eta_t >= 1.0:
            continue
        inc = incoming(th, inside)
        n = 2000
        uv = rng.random((n, 2)).astype(np.float32)
        out, filt, pdf = psyref.closure_sample(K, inc, N, N, uv, color=WHITE, roughness=1e-4, ior=ior,
                                               w=np.full(n, 0.9999, np.float32))
        ok = pdf > 0
        assert ok.mean() > 0.99
        o = out[ok].astype(np.float64)
        o /= np.linalg.norm(o, axis=1, keepdims=True)
        # transmitted to the other side, in the plane of incidence
        assert np.all(np.sign(o[:, 2]) == np.sign(inc[2]))
        assert np.max(np.abs(o[:, 1])) < 2e-3
        sin_t = np.hypot(o[:, 0], o[:, 1])
        assert np.allclose(eta_t * sin_t, eta_i * math.sin(th), atol=2e-3), theta_deg


def uniform_sphere(n, rng):
    z = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)


@pytest.mark.parametrize("inside", [False, True])
@pytest.mark.parametrize("alpha", [0.15, 0.4])
def test_density_integrates_to_the_sampled_fraction(alpha, inside):
    ior = 1.5
    rng = np.random.default_rng(11)
    inc = incoming(math.radians(35.0), inside)
    n = 400_000
    dirs = uniform_sphere(n, rng)
    _, pdf = psyref.closure_evaluate(K, inc, dirs, N, N, color=WHITE, roughness=alpha, ior=ior)
    integral = float(np.mean(pdf.astype(np.float64))) * 4 * math.pi
    refl_side = (dirs[:, 2] > 0) == (inc[2] < 0)
    refl_integral = float(np.mean(np.where(refl_side, pdf, 0).astype(np.float64))) * 4 * math.pi
    m = 200_000
    uvw = rng.random((m, 3)).astype(np.float32)
    out, _, spdf = psyref.closure_sample(K, inc, N, N, uvw[:, :2], color=WHITE, roughness=alpha, ior=ior,
                                         w=uvw[:, 2])
    valid = spdf > 0
    reflected = valid & ((out[:, 2] > 0) == (inc[2] < 0))
    # Monte Carlo of the integral: a few parts in a thousand at these sizes.
    assert integral == pytest.approx(valid.mean(), abs=0.02), (integral, valid.mean())
    assert refl_integral == pytest.approx(reflected.mean(), abs=0.02), (refl_integral, reflected.mean())
    assert valid.mean() > 0.75


@pytest.mark.parametrize("inside", [False, True])
def test_white_furnace(inside):
    """The mean sample weight of a white dielectric: at most 1 (no energy
    created), within 1% of 1 when nearly smooth at every angle, and falling
    as the roughness grows (single-scattering microfacets lose the light that
    would bounce between facets)."""
    rng = np.random.default_rng(13)
    for theta_deg in (0.0, 30.0, 60.0, 75.0):
        inc = incoming(math.radians(theta_deg), inside)
        means = []
        for alpha in (0.02, 0.1, 0.3, 0.8):
            m = 200_000
            uvw = rng.random((m, 3)).astype(np.float32)
            _, filt, pdf = psyref.closure_sample(K, inc, N, N, uvw[:, :2], color=WHITE, roughness=alpha, ior=1.5,
                                                 w=uvw[:, 2])
            wgt = np.where(pdf > 0, filt[:, 0].astype(np.float64) / np.maximum(pdf, 1e-30), 0.0)
            means.append(float(wgt.mean()))
        assert max(means) <= 1.0 + 0.005, (theta_deg, inside, means)
        assert means[0] >= 0.99, (theta_deg, inside, means)
        assert all(a >= b - 0.005 for a, b in zip(means, means[1:])), (theta_deg, inside, means)


def test_reflection_is_reciprocal():
    rng = np.random.default_rng(17)
    for _ in range(200):
        a = uniform_sphere(1, rng)[0]
        b = uniform_sphere(1, rng)[0]
        a[2], b[2] = abs(a[2]) + 0.05, abs(b[2]) + 0.05
        fa, _ = psyref.closure_evaluate(K, -a, b, N, N, color=WHITE, roughness=0.3, ior=1.5)
        fb, _ = psyref.closure_evaluate(K, -b, a, N, N, color=WHITE, roughness=0.3, ior=1.5)
        a_n, b_n = a / np.linalg.norm(a), b / np.linalg.norm(b)
        # filter = f |cos(out)|
        assert fa[0, 0] / b_n[2] == pytest.approx(fb[0, 0] / a_n[2], rel=1e-4, abs=1e-7)


def test_transmission_is_tinted_and_reflection_is_not():
    inc = incoming(math.radians(30.0))
    tint = (0, 0.2, 0.5, 0.9)
    refr = np.array([0.25, 0.0, -0.968], np.float32)
    refl = np.array([0.5, 0.0, 0.866], np.float32)
    ft, _ = psyref.closure_evaluate(K, inc, refr, N, N, color=tint, roughness=0.3, ior=1.5)
    fw, _ = psyref.closure_evaluate(K, inc, refr, N, N, color=WHITE, roughness=0.3, ior=1.5)
    rt, _ = psyref.closure_evaluate(K, inc, refl, N, N, color=tint, roughness=0.3, ior=1.5)
    rw, _ = psyref.closure_evaluate(K, inc, refl, N, N, color=WHITE, roughness=0.3, ior=1.5)
    assert ft[0, 0] > 0 and not np.allclose(ft, fw)
    assert np.array_equal(rt, rw) and rt[0, 0] > 0
