"""Pins the oracle's GGX, rough dielectric and Disney BXDFs with what the mathematics fixes,
chosen so that each plausible transcription slip fails one check:

GGX (scatteringevent/bsdf/microfacet.cpp:229-284, 320-356; microfacet.h:95-160):
* D is a normalised microfacet density: the integral of D(h) cos(theta_h) over the hemisphere is
  1. D is read back from the oracle's pdf: with wo = wi = h, pdf = D(h) cos(theta_h) / (4 |wo.h|)
  = D(h) cos(theta_h) / 4. A D scaled by any factor, or a wrong alpha in one place, fails.
* The reflection pdf at normal incidence integrates to 1 / (1 + alpha^2) exactly (the mass of
  D cos beyond 45 degrees, whose mirror directions fall below the horizon, is alpha^2/(1+alpha^2)
  for GGX). This pins the 1/(4|wo.h|) Jacobian and alpha = roughness^2.
* Smith masking obeys the projected-area identity (Heitz 2014, eq. 7): for every wo,
  G1(wo) * integral of max(0, wo.h) D(h) dw_h = cos(theta_o). G1(wo)^2 is read back from f/pdf at
  wi = wo with a Fresnel of 1 - 4e-6 (conductor with k = 1000).

Rough dielectric (dielectric.cpp:32-66, fresnel.h:53-73, microfacet.cpp:390-440):
* With white reflectance and transmittance, sample_f picks reflection with probability exactly
  F(cos theta_o), the unpolarised Fresnel reflectance of a 1 | 1.5 interface (Snell's law, with
  total internal reflection from inside beyond asin(1/1.5)). Swapped indices or a swapped
  branch fail at oblique angles.
* Smooth-limit furnace: a closed dielectric sphere (roughness 0.02) under a constant environment
  renders the environment's radiance. A transmission lobe without its (1 - F), or a Jacobian
  error, gains or loses energy at every interface.

Disney (disney.cpp:181-434), each lobe isolated by its parameters:
* Diffuse + retro-reflection (Burley 2015, eq. 4) at wo = up: the hemispherical albedo is
  base * (41/42 + 5 r / 84) (Beta-function integrals of the Schlick weights), and f at
  wo = wi = up is base / pi.
* Sheen alone (base 0): albedo = sheen * 2 pi * integral of (1 - cos(theta/2))^5 cos(theta).
* Metallic 1 at roughness -> 0: E[f/pdf] is Schlick's mirror reflectance base + (1-base)(1-cos)^5.
* Specular and clearcoat lobes alone: their microfacet densities (GGX and GTR1) integrate to 1.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O
from sort_amd import integrator as I
from sort_amd import scenes

F32 = np.float32


def local_dirs(theta, phi=0.3):
    theta = np.asarray(theta, np.float64)
    return np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], 1).astype(F32)


def theta_grid(n=200000, top=math.pi / 2):
    """Midpoint rule nodes and weights for the integral over the hemisphere of a function of
    theta alone: 2 pi * sum g(theta) sin(theta) dtheta."""
    d = top / n
    th = (np.arange(n) + 0.5) * d
    return th, 2 * math.pi * np.sin(th) * d


def density_integral(mat):
    """Integral over the hemisphere of D(h) cos(theta_h), with D cos = 4 pdf(h, h)."""
    th, w = theta_grid()
    h = local_dirs(th)
    dcos = 4.0 * O.bxdf_pdf(mat, h, h).astype(np.float64)
    return float(np.sum(dcos * w))


# ---- GGX --------------------------------------------------------------------------------------

@pytest.mark.parametrize("rough", [0.2, 0.45, 0.8])
def test_ggx_density_is_normalised(rough):
    assert abs(density_integral(I.ggx((1, 1, 1), rough)) - 1.0) < 2e-3


@pytest.mark.parametrize("rough", [0.3, 0.6, 0.9])
def test_ggx_reflection_pdf_mass_at_normal_incidence(rough):
    alpha = max(rough, 1e-3) ** 2
    th, w = theta_grid()
    wi = local_dirs(th)
    wo = np.tile(np.array([[0, 1, 0]], F32), (len(th), 1))
    mass = float(np.sum(O.bxdf_pdf(I.ggx((1, 1, 1), rough), wo, wi).astype(np.float64) * w))
    assert abs(mass - 1.0 / (1.0 + alpha * alpha)) < 2e-3, (mass, 1 / (1 + alpha * alpha))


def _phi_integral_pos(a, b):
    """Integral over phi in [0, 2pi) of max(0, a cos(phi) + b), a >= 0."""
    out = np.where(b >= a, 2 * math.pi * b, 0.0)
    m = (b < a) & (b > -a)
    phi0 = np.arccos(np.clip(-b[m] / np.maximum(a[m], 1e-300), -1, 1))
    out[m] = 2 * (b[m] * phi0 + a[m] * np.sin(phi0))
    return out


@pytest.mark.parametrize("rough", [0.35, 0.7])
@pytest.mark.parametrize("theta_o_deg", [20.0, 50.0, 75.0])
def test_ggx_smith_masking_projected_area_identity(rough, theta_o_deg):
    m = I.ggx((1, 1, 1), rough, eta=(1.0, 1.0, 1.0), k=(1000.0, 1000.0, 1000.0))
    tho = math.radians(theta_o_deg)
    wo = local_dirs([tho], phi=0.0)
    f = O.bxdf_f(m, wo, wo)[0, 0]
    pdf = O.bxdf_pdf(m, wo, wo)[0]
    g1 = math.sqrt(float(f) / float(pdf) * math.cos(tho) ** 2 / (1 - 4e-6))
    th, _ = theta_grid(200000)
    dth = th[1] - th[0]
    h = local_dirs(th)
    d = 4.0 * O.bxdf_pdf(m, h, h).astype(np.float64) / np.cos(th)
    proj = _phi_integral_pos(math.sin(tho) * np.sin(th), math.cos(tho) * np.cos(th))
    lhs = g1 * float(np.sum(d * proj * np.sin(th) * dth))
    assert abs(lhs - math.cos(tho)) < 2e-3 * math.cos(tho) + 2e-4, (lhs, math.cos(tho), g1)


# ---- rough dielectric -------------------------------------------------------------------------

def fresnel_unpolarised(cos_i, n1, n2):
    sin_t = n1 / n2 * math.sqrt(max(0.0, 1 - cos_i * cos_i))
    if sin_t >= 1:
        return 1.0
    cos_t = math.sqrt(1 - sin_t * sin_t)
    rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
    rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)
    return 0.5 * (rs * rs + rp * rp)


@pytest.mark.parametrize("theta_deg,inside", [(0.0, False), (45.0, False), (70.0, False), (85.0, False),
                                              (20.0, True), (38.0, True), (50.0, True)])
def test_dielectric_reflect_fraction_is_fresnel(theta_deg, inside):
    m = I.dielectric((1, 1, 1), (1, 1, 1), roughness=0.05, ior=1.5)
    th = math.radians(theta_deg)
    wo = local_dirs([th], phi=0.0)
    if inside:
        wo[:, 1] = -wo[:, 1]
    n = 1 << 17
    g = np.random.default_rng(7)
    _, wi, pdf = O.bxdf_sample(m, np.repeat(wo, n, 0), g.random((n, 7)).astype(F32))
    same = np.sign(wi[:, 1]) == np.sign(wo[0, 1])
    frac = float(same.mean())
    want = fresnel_unpolarised(math.cos(th), 1.5, 1.0) if inside else fresnel_unpolarised(math.cos(th), 1.0, 1.5)
    sd = math.sqrt(max(want * (1 - want), 1e-6) / n)
    assert abs(frac - want) < 5 * sd + 2e-3, (frac, want)


def test_dielectric_smooth_furnace_sphere():
    env = (1.0, 0.5, 0.25)
    tris = scenes.uv_sphere((0.0, 0.0, 0.0), 0.5, 128, 64)
    pos, idx = scenes.tris_to_mesh(tris)
    nrm = (pos / np.linalg.norm(pos, axis=1, keepdims=True)).astype(F32)
    sc = {"pos": pos, "idx": idx, "quads": [], "normals": nrm, "tri_material": np.zeros(len(idx), np.uint32),
          "materials": [I.dielectric((1, 1, 1), (1, 1, 1), roughness=0.02, ior=1.5)],
          "quad_radiance": np.zeros((0, 3), F32), "env": env}
    pt = O.OraclePT(sc)
    W = H = 24
    cam = dict(eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_y_deg=22.0)
    img = pt.render(cam, W, H, 128, max_depth=8, seed=5, threads=8).reshape(H, W, 3)
    ys = 1.0 / math.tan(math.radians(22.0) / 2)
    hit = []
    for y in range(H):
        for x in range(W):
            d = np.array([(2 * (x + 0.5) / W - 1) / ys, (1 - 2 * (y + 0.5) / H) / ys, -1.0])
            d /= np.linalg.norm(d)
            o = np.array([0.0, 0.0, 3.0])
            b = np.dot(o, d)
            if b * b - (np.dot(o, o) - 0.25) > 0.05:    # clearly on the sphere, away from the rim
                hit.append(img[y, x])
    hv = np.array(hit)
    assert len(hv) > 100
    assert np.allclose(hv.mean(axis=0), env, rtol=0.02), hv.mean(axis=0)


# ---- Disney -----------------------------------------------------------------------------------

def albedo_up(mat):
    """Integral over the upper hemisphere of f(up, wi) (SORT's f carries |cos wi|)."""
    th, w = theta_grid(100000)
    wi = local_dirs(th)
    wo = np.tile(np.array([[0, 1, 0]], F32), (len(th), 1))
    return (O.bxdf_f(mat, wo, wi).astype(np.float64) * w[:, None]).sum(axis=0)


@pytest.mark.parametrize("rough", [0.0, 0.5, 1.0])
def test_disney_diffuse_and_retro_reflection_albedo(rough):
    base = np.array([0.8, 0.5, 0.2])
    m = I.disney(tuple(base), metallic=0.0, specular=0.0, roughness=rough, sheen=0.0, clearcoat=0.0)
    got = albedo_up(m)
    want = base * (41.0 / 42.0 + 5.0 * rough / 84.0)
    assert np.allclose(got, want, rtol=1e-3), (got, want)
    up = np.array([[0, 1, 0]], F32)
    assert np.allclose(O.bxdf_f(m, up, up)[0], base / math.pi, rtol=1e-6)


@pytest.mark.parametrize("sheen", [0.4, 1.0])
def test_disney_sheen_albedo(sheen):
    m = I.disney((0.0, 0.0, 0.0), metallic=0.0, specular=0.0, roughness=0.5, sheen=sheen, sheen_tint=0.7,
                 clearcoat=0.0)
    got = albedo_up(m)
    val, _ = integrate.quad(lambda t: (1 - math.cos(t / 2)) ** 5 * math.cos(t) * math.sin(t), 0, math.pi / 2,
                            epsabs=1e-13)
    want = sheen * 2 * math.pi * val
    assert np.allclose(got, want, rtol=2e-3), (got, want)


@pytest.mark.parametrize("theta_deg", [0.0, 40.0, 70.0, 82.0])
def test_disney_metal_mirror_limit_is_schlick(theta_deg):
    base = np.array([0.9, 0.6, 0.3])
    m = I.disney(tuple(base), metallic=1.0, specular=0.5, roughness=0.02, clearcoat=0.0, sheen=0.0)
    th = math.radians(theta_deg)
    wo = local_dirs([th], phi=0.0)
    n = 1 << 14
    g = np.random.default_rng(9)
    f, _, pdf = O.bxdf_sample(m, np.repeat(wo, n, 0), g.random((n, 7)).astype(F32))
    ok = pdf > 0
    est = np.where(ok[:, None], f / np.where(ok, pdf, 1)[:, None], 0).mean(axis=0)
    want = base + (1 - base) * (1 - math.cos(th)) ** 5
    assert np.allclose(est, want, rtol=0.01), (est, want)


@pytest.mark.parametrize("rough", [0.3, 0.7])
def test_disney_specular_density_is_normalised(rough):
    m = I.disney((0.0, 0.0, 0.0), metallic=0.0, specular=1.0, roughness=rough, sheen=0.0, clearcoat=0.0)
    assert abs(density_integral(m) - 1.0) < 2e-3


@pytest.mark.parametrize("gloss", [0.0, 0.5, 0.8])
def test_disney_clearcoat_density_is_normalised(gloss):
    m = I.disney((0.0, 0.0, 0.0), metallic=0.0, specular=0.0, roughness=0.5, sheen=0.0, clearcoat=1.0,
                 clearcoat_gloss=gloss)
    assert abs(density_integral(m) - 1.0) < 2e-3
