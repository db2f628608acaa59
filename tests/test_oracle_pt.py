"""Pins the oracle's path tracer without the reference binary:

* the BXDF statistical suite of unittests/bxdf.cpp:41-120 (pdf of Sample_F vs Pdf within 1%,
  f of Sample_F vs F within 1e-3, integral of the pdf <= 1.03, E[1/pdf] = 2 pi for cosine
  sampling, reciprocity, energy <= 1.03), with the enabled / disabled split of bxdf.cpp:131-195;
* a furnace: a Lambert sphere (albedo rho) under a constant environment renders rho * L_env;
* the closed-form irradiance of a quad light (Lambert's polygon formula) on a diffuse plane,
  which checks both halves of the MIS next-event estimator.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from sort_amd import integrator as I
from sort_amd import scenes


def _rng(seed):
    return np.random.default_rng(seed)


def _sphere_dirs(g, n):
    v = g.normal(size=(n, 3)).astype(np.float32)
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def _hemi_dirs(g, n):
    v = _sphere_dirs(g, n)
    v[:, 1] = np.abs(v[:, 1])
    return v


MATS = {
    "lambert": I.lambert((0.9, 0.6, 0.3)),
    "ggx": I.ggx((1.0, 1.0, 1.0), 0.35),
    "ggx_rough": I.ggx((1.0, 1.0, 1.0), 0.8),
    "dielectric": I.dielectric((1, 1, 1), (1, 1, 1), roughness=0.3, ior=1.5),
    "disney": I.disney((0.8, 0.5, 0.3), metallic=0.3, specular=0.5, specular_tint=0.2, roughness=0.4, sheen=0.3,
                       clearcoat=0.5, clearcoat_gloss=0.7),
}


@pytest.mark.parametrize("name", ["lambert", "ggx", "ggx_rough", "disney"])
def test_sample_f_matches_f_and_pdf(name):
    """bxdf.cpp:79-96 (checkPdf, first block)."""
    m = MATS[name]
    g = _rng(1)
    n = 4096
    wo = np.repeat(_hemi_dirs(g, 1), n, axis=0)
    u = g.random((n, 7)).astype(np.float32)
    f0, wi, pdf = O.bxdf_sample(m, wo, u)
    pdf1 = O.bxdf_pdf(m, wo, wi)
    f1 = O.bxdf_f(m, wo, wi)
    ok = pdf1 > 0
    assert ok.mean() > 0.5
    assert np.all(np.abs(pdf[ok] / pdf1[ok] - 1.0) <= 0.01)
    assert np.all(np.abs(f0[ok] - f1[ok]) <= 1e-3)


@pytest.mark.parametrize("name", ["lambert", "ggx", "ggx_rough", "dielectric"])
def test_pdf_integrates_to_at_most_one(name):
    """bxdf.cpp:101-108: E_uniform-sphere[pdf / (1/4pi)] <= 1.03. Disney is excluded as in the
    reference (DISABLED_Disney, bxdf.cpp:163-171): its diffuse pdf term CosHemispherePdf has no
    hemisphere test (disney.cpp:423-427), so its pdf integrates above 1 (checked below)."""
    m = MATS[name]
    g = _rng(2)
    n = 1 << 20
    wo = np.repeat(_hemi_dirs(g, 1), n, axis=0)
    wi = _sphere_dirs(g, n)
    total = float(np.mean(O.bxdf_pdf(m, wo, wi) / (1.0 / (4 * math.pi))))
    assert total <= 1.03
    if name == "lambert":
        assert abs(total - 1.0) < 0.01  # a full cosine lobe above a wo in the upper hemisphere


def test_disney_pdf_mass_matches_its_lobe_weights():
    """The restated Disney pdf = sum_lobe w_lobe * pdf_lobe / sum w with a two-sided cosine term,
    so its integral is (w_cc * I_cc + w_sr * I_sr + 2 * w_dr) / total with I_cc, I_sr <= 1."""
    m = MATS["disney"]
    g = _rng(2)
    n = 1 << 20
    wo = np.repeat(_hemi_dirs(g, 1), n, axis=0)
    wi = _sphere_dirs(g, n)
    total = float(np.mean(O.bxdf_pdf(m, wo, wi) / (1.0 / (4 * math.pi))))
    lum = lambda c: 0.212671 * c[0] + 0.715160 * c[1] + 0.072169 * c[2]
    base = np.array(m["base"])
    ctint = base / lum(base)
    cspec0 = (m["specular"] * 0.04 * ((1 - m["specular_tint"]) + m["specular_tint"] * ctint)) * (1 - m["metallic"]) \
        + base * m["metallic"]
    w_cc, w_sr = m["clearcoat"] * 0.04, lum(cspec0) * 8 * (1 - m["roughness"])
    w_dr = lum(base) * (1 - m["metallic"])
    hi = (w_cc + w_sr + 2 * w_dr) / (w_cc + w_sr + w_dr)
    lo = 2 * w_dr / (w_cc + w_sr + w_dr)
    assert lo - 0.02 <= total <= hi + 0.02, (lo, total, hi)


def test_lambert_inverse_pdf_expectation_is_two_pi():
    """bxdf.cpp:111-119: E[1/pdf] over the sampled directions = the sampled solid angle."""
    g = _rng(3)
    n = 1 << 20
    wo = np.repeat(_hemi_dirs(g, 1), n, axis=0)
    _, _, pdf = O.bxdf_sample(MATS["lambert"], wo, g.random((n, 7)).astype(np.float32))
    ok = pdf != 0
    total = float(np.sum(1.0 / pdf[ok]) / n)
    assert abs(total - 2 * math.pi) < 0.03


@pytest.mark.parametrize("name", ["lambert", "ggx", "ggx_rough"])
def test_reciprocity(name):
    """bxdf.cpp:42-57: F(wo,wi)|cos wo| == F(wi,wo)|cos wi| (SORT's F carries |cos wi|)."""
    m = MATS[name]
    g = _rng(4)
    n = 1024
    wo, wi = _sphere_dirs(g, n), _sphere_dirs(g, n)
    f0 = O.bxdf_f(m, wo, wi) * np.abs(wo[:, 1:2])
    f1 = O.bxdf_f(m, wi, wo) * np.abs(wi[:, 1:2])
    assert np.all(np.abs(f0 - f1) <= 1e-3)


@pytest.mark.parametrize("name", ["lambert", "ggx", "ggx_rough", "dielectric"])
def test_energy_conservation(name):
    """bxdf.cpp:59-69: E[f/pdf] for wo = DIR_UP stays <= 1.03."""
    m = MATS[name]
    g = _rng(5)
    n = 1 << 20
    wo = np.tile(np.array([[0.0, 1.0, 0.0]], np.float32), (n, 1))
    f, _, pdf = O.bxdf_sample(m, wo, g.random((n, 7)).astype(np.float32))
    ok = pdf > 0
    est = np.zeros_like(f)
    est[ok] = f[ok] / pdf[ok, None]
    total = est.mean(axis=0)
    assert np.all(total <= 1.03), total
    if name == "lambert":
        assert np.allclose(total, [0.9, 0.6, 0.3], rtol=0.01)  # white-furnace albedo


def test_ggx_conductor_fresnel_normal_incidence():
    """fresnel.h ConductorFresnel at cos = 1 is ((eta-1)^2 + k^2) / ((eta+1)^2 + k^2)."""
    eta, k = np.array([0.27, 0.68, 1.22]), np.array([3.6, 2.6, 2.3])
    want = ((eta - 1) ** 2 + k ** 2) / ((eta + 1) ** 2 + k ** 2)
    m = I.ggx((1, 1, 1), 0.001, eta=tuple(eta), k=tuple(k))
    # for a near-mirror, E[f/pdf] at normal incidence approaches F(1) * G ~ F(1)
    g = _rng(6)
    n = 1 << 16
    wo = np.tile(np.array([[0.0, 1.0, 0.0]], np.float32), (n, 1))
    f, _, pdf = O.bxdf_sample(m, wo, g.random((n, 7)).astype(np.float32))
    ok = pdf > 0
    est = (f[ok] / pdf[ok, None]).mean(axis=0)
    assert np.allclose(est, want, rtol=0.02), (est, want)


def _sphere_in_env(rho, env, segments=128, rings=64):
    tris = scenes.uv_sphere((0.0, 0.0, 0.0), 0.5, segments, rings)
    pos, idx = scenes.tris_to_mesh(tris)
    nrm = (pos / np.linalg.norm(pos, axis=1, keepdims=True)).astype(np.float32)
    return {"pos": pos, "idx": idx, "quads": [], "normals": nrm, "tri_material": np.zeros(len(idx), np.uint32),
            "materials": [I.lambert(rho)], "quad_radiance": np.zeros((0, 3), np.float32), "env": env}


def test_furnace_lambert_sphere_under_constant_environment():
    rho, env = (0.6, 0.3, 0.9), (1.0, 0.5, 0.25)
    sc = _sphere_in_env(rho, env)
    pt = O.OraclePT(sc)
    W = H = 32
    cam = dict(eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_y_deg=25.0)
    img = pt.render(cam, W, H, 256, max_depth=4, seed=3, threads=8).reshape(H, W, 3)
    # pixel centres that clearly hit / miss the sphere (analytic ray-sphere test, margin 3%)
    ys = 1.0 / math.tan(math.radians(25.0) / 2)
    hit, miss = [], []
    for y in range(H):
        for x in range(W):
            cx = 2 * (x + 0.5) / W - 1
            cy = 1 - 2 * (y + 0.5) / H
            d = np.array([cx / ys, cy / ys, -1.0])
            d /= np.linalg.norm(d)
            o = np.array([0.0, 0.0, 3.0])
            b = np.dot(o, d)
            disc = b * b - (np.dot(o, o) - 0.25)
            (hit if disc > 0.03 else miss if disc < -0.03 else []).append((y, x))
    hv = np.array([img[y, x] for y, x in hit])
    mv = np.array([img[y, x] for y, x in miss])
    assert len(hv) > 50 and len(mv) > 50
    assert np.allclose(mv, env, rtol=1e-6)
    assert np.allclose(hv.mean(axis=0), np.array(rho) * np.array(env), rtol=0.01), hv.mean(axis=0)


def _polygon_irradiance(p, n, verts, L):
    """Lambert's formula: E = L/2 * |sum_i theta_i (gamma_i . n)| for a Lambertian polygon."""
    u = [(v - p) / np.linalg.norm(v - p) for v in verts]
    s = 0.0
    for i in range(len(u)):
        a, b = u[i], u[(i + 1) % len(u)]
        th = math.acos(np.clip(np.dot(a, b), -1, 1))
        c = np.cross(a, b)
        s += th * np.dot(c / np.linalg.norm(c), n)
    return L / 2 * abs(s)


def test_quad_light_direct_irradiance_closed_form():
    rho, L = 0.5, 10.0
    ground = [np.array([[-10, 0, -10], [10, 0, 10], [10, 0, -10]], np.float32),
              np.array([[-10, 0, -10], [-10, 0, 10], [10, 0, 10]], np.float32)]
    pos, idx = scenes.tris_to_mesh(ground)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (len(pos), 1))
    light = scenes.translate_quad((0.0, 1.0, 0.0), 1.0, facing_down=True)
    sc = {"pos": pos, "idx": idx, "quads": [light], "normals": nrm, "tri_material": np.zeros(2, np.uint32),
          "materials": [I.lambert((rho, rho, rho))], "quad_radiance": np.array([[L, L, L]], np.float32),
          "env": (0, 0, 0)}
    pt = O.OraclePT(sc)
    pts = [(0.0, 0.0), (0.3, 0.1), (0.7, -0.4), (1.5, 0.0), (2.5, 2.0)]
    spp = 1 << 14
    rays, keys = [], []
    for i, (x, z) in enumerate(pts):
        for k in range(spp):
            rays.append([x, 0.5, z, 0.0, 0.0, -1.0, 0.0, 3.4028234663852886e38])
            keys.append([i, k])
    out = pt.li(np.array(rays, np.float32), np.array(keys, np.uint32), max_depth=1, seed=11, threads=8)
    est = out.reshape(len(pts), spp, 3).mean(axis=1)[:, 0]
    corners = [np.array(c) for c in [(-0.5, 1, -0.5), (0.5, 1, -0.5), (0.5, 1, 0.5), (-0.5, 1, 0.5)]]
    for (x, z), e in zip(pts, est):
        E = _polygon_irradiance(np.array([x, 0.0, z]), np.array([0, 1, 0]), corners, L)
        assert abs(e - rho / math.pi * E) <= 0.01 * rho / math.pi * E + 1e-4, (x, z, e, rho / math.pi * E)


def test_ray_dump_is_consistent_with_the_traversal_and_the_camera():
    """The oracle's ray dump (every extension ray of PathTracing::li with its closest hit): bounce 0
    is the camera ray of the pixel's jittered raster point, each recorded hit is what the BVH gives
    for the recorded ray, a bounce is traced only after the one before it, and the number of
    recorded rays equals the render's extension-ray count."""
    sc = scenes.cornell_sphere()
    pt = O.OraclePT(sc)
    W = H = 16
    spp, depth = 4, 5
    rays, hits, valid = pt.ray_dump(sc["camera"], W, H, spp, depth, 3, 0, W * H, threads=4)
    st = {}
    pt.render(sc["camera"], W, H, spp, depth, 3, threads=4, stats=st)
    assert valid.sum() == st["extension_rays"]
    v = valid.reshape(W * H, spp, depth)
    assert v[:, :, 0].all()
    assert not np.any(v[:, :, 1:] & ~v[:, :, :-1])
    again = pt.bvh.closest(rays[valid], threads=4)
    assert np.array_equal(again.view(np.uint32), hits[valid].view(np.uint32))
    # camera rays start at the eye with tmin 0 and point into the room (+z)
    r0 = rays.reshape(W * H, spp, depth, 8)[:, :, 0]
    assert np.allclose(r0[..., 0:3], sc["camera"]["eye"]) and np.all(r0[..., 3] == 0.0)
    assert np.all(r0[..., 6] > 0.0)
    # later bounces start where the previous ray hit, with tmin 1e-4
    r1 = rays.reshape(W * H, spp, depth, 8)[:, :, 1]
    h0 = hits.reshape(W * H, spp, depth, 4)[:, :, 0]
    m = v[:, :, 1]
    p = r0[..., 0:3] + h0[..., 0:1] * r0[..., 4:7]
    assert np.allclose(r1[m][:, 0:3], p[m], atol=1e-5) and np.all(r1[m][:, 3] == np.float32(1e-4))


def test_per_sample_clamp_closed_forms():
    """ImageEvaluation's clamping (image_evaluation.cpp:222-226): every sample's radiance is clamped
    to [0, clamp] before the mean. Looking straight at the 17-radiance light from below, every
    sample carries the light's emission (plus what its white surface reflects), so with a clamp of
    5 every sample is 5 and so is the pixel; a clamp above every sample changes no bit; a clamped
    image never exceeds the clamp."""
    sc = scenes.cornell_sphere()
    cam = dict(eye=(0.5, 0.9, 0.5), target=(0.5, 1.0, 0.5), up=(0.0, 0.0, 1.0), fov_y_deg=5.0)
    plain = O.OraclePT(sc).render(cam, 8, 8, 4, 5, 3, threads=4)
    assert np.all(plain >= np.float32(17.0))
    clamped = O.OraclePT(dict(sc, clamp=5.0)).render(cam, 8, 8, 4, 5, 3, threads=4)
    assert np.all(clamped == np.float32(5.0))
    room = sc["camera"]
    a = O.OraclePT(sc).render(room, 24, 24, 4, 5, 9, threads=8)
    b = O.OraclePT(dict(sc, clamp=1e30)).render(room, 24, 24, 4, 5, 9, threads=8)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    c = O.OraclePT(dict(sc, clamp=0.5)).render(room, 24, 24, 4, 5, 9, threads=8)
    assert np.all(c <= np.float32(0.5)) and np.all(c >= 0) and np.any(a > 0.5)


def test_transparent_sheet_attenuates_direct_light_closed_form():
    """SORT's transparent shadows (ENABLE_TRANSPARENT_SHADOW, integratormethod.cpp:121-145 and
    168-186; Scene::GetAttenuation, core/scene.cpp:192-208) and its Transparent BXDF
    (bsdf/transparent.cpp): a sheet with attenuation A between a Lambert ground and a quad light
    scales the ground's direct light by A in both MIS halves; seen through the sheet (the camera
    ray crosses it first, continuing straight with weight A), the ground's direct light is scaled
    by A twice. Expected values from Lambert's polygon irradiance formula."""
    rho, L = 0.5, 10.0
    A = np.array([0.5, 0.25, 0.8])
    ground = [np.array([[-10, 0, -10], [10, 0, 10], [10, 0, -10]], np.float32),
              np.array([[-10, 0, -10], [-10, 0, 10], [10, 0, 10]], np.float32)]
    sheet = [np.array([[-10, 0.7, -10], [10, 0.7, 10], [10, 0.7, -10]], np.float32),
             np.array([[-10, 0.7, -10], [-10, 0.7, 10], [10, 0.7, 10]], np.float32)]
    pos, idx = scenes.tris_to_mesh(ground + sheet)
    nrm = np.tile(np.array([[0, 1, 0]], np.float32), (len(pos), 1))
    light = scenes.translate_quad((0.0, 1.0, 0.0), 1.0, facing_down=True)
    sc = {"pos": pos, "idx": idx, "quads": [light], "normals": nrm,
          "tri_material": np.array([0, 0, 1, 1], np.uint32),
          "materials": [I.lambert((rho, rho, rho)), I.transparent(tuple(A))],
          "quad_radiance": np.array([[L, L, L]], np.float32), "env": (0, 0, 0)}
    pt = O.OraclePT(sc)
    pts = [(0.0, 0.0), (0.3, 0.1), (0.7, -0.4), (1.5, 0.0)]
    spp = 1 << 14
    corners = [np.array(c) for c in [(-0.5, 1, -0.5), (0.5, 1, -0.5), (0.5, 1, 0.5), (-0.5, 1, 0.5)]]
    for y0, depth, power in ((0.5, 1, 1), (0.9, 2, 2)):  # below the sheet; above it, looking through
        rays, keys = [], []
        for i, (x, z) in enumerate(pts):
            for k in range(spp):
                rays.append([x, y0, z, 0.0, 0.0, -1.0, 0.0, 3.4028234663852886e38])
                keys.append([i, k])
        out = pt.li(np.array(rays, np.float32), np.array(keys, np.uint32), max_depth=depth, seed=13, threads=8)
        est = out.reshape(len(pts), spp, 3).mean(axis=1)
        for (x, z), e in zip(pts, est):
            E = _polygon_irradiance(np.array([x, 0.0, z]), np.array([0, 1, 0]), corners, L)
            want = A ** power * rho / math.pi * E
            assert np.all(np.abs(e - want) <= 0.01 * want + 1e-4), (y0, x, z, e, want)
