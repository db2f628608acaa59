"""Transparent materials in the path tracer (SORT's Transparent BXDF and its transparent shadows,
ENABLE_TRANSPARENT_SHADOW: EvaluateDirect's Visibility::GetAttenuation, integratormethod.cpp:121-145
and 168-186, over Scene::GetAttenuation, core/scene.cpp:192-208): the GPU render of a Cornell box
with tinted transparent panels (one across the light, one in front of the sphere) against the
oracle with the same random-number keys, the batched Li form, and the served form with the
transparent scene among opaque ones (each result bitwise its single-request render)."""
import numpy as np
import pytest

from oracle.oracle import OraclePT
from sort_amd import Obvh, integrator as I, scenes
from sort_amd import serve as S

pytestmark = pytest.mark.gpu


def panel_scene():
    sc = scenes.cornell_sphere()
    A = (0.6, 0.9, 0.4)
    panels = [scenes._quad_tris(np.array(p0, np.float32), np.array(p1, np.float32), np.array(p2, np.float32),
                                np.array(p3, np.float32)) for p0, p1, p2, p3 in [
        # horizontal sheet just under the light: every light-sampled shadow ray crosses it
        ((0.3, 0.9, 0.3), (0.7, 0.9, 0.3), (0.7, 0.9, 0.7), (0.3, 0.9, 0.7)),
        # vertical pane between the camera and the sphere
        ((0.2, 0.15, 0.2), (0.65, 0.15, 0.2), (0.65, 0.75, 0.2), (0.2, 0.75, 0.2)),
    ]]
    tris = [t for p in panels for t in p]
    extra_pos, extra_idx = scenes.tris_to_mesh(tris)
    n0 = len(sc["pos"])
    sc["pos"] = np.concatenate([sc["pos"], extra_pos]).astype(np.float32)
    sc["idx"] = np.concatenate([sc["idx"], extra_idx + n0]).astype(np.uint32)
    sc["normals"] = np.concatenate([sc["normals"], scenes.face_normals_per_vertex(extra_pos, extra_idx)]).astype(
        np.float32)
    sc["materials"] = sc["materials"] + [I.transparent(A)]
    sc["tri_material"] = np.concatenate([sc["tri_material"], np.full(len(extra_idx), len(sc["materials"]) - 1,
                                                                       np.uint32)])
    sc["ntris"] = len(sc["idx"])
    return sc


def _lum(img):
    return 0.212671 * img[..., 0] + 0.715160 * img[..., 1] + 0.072169 * img[..., 2]


def _pt(ctx, sc, depth=5):
    acc = Obvh(ctx).build(sc["pos"], sc["idx"], sc["quads"])
    return acc, I.PathTracing(acc, sc["materials"], tri_material=sc["tri_material"], normals=sc["normals"],
                              quad_radiance=sc["quad_radiance"], env_radiance=sc["env"], max_recursive_depth=depth)


def test_transparent_panels_same_keys_match_oracle(gpu_ctx):
    sc = panel_scene()
    acc, pt = _pt(gpu_ctx, sc)
    W, H, spp = 96, 96, 16
    print(f"[transp] gpu {W}x{H} {spp} spp", flush=True)
    g = pt.render(I.camera(**sc["camera"]), W, H, spp, seed=21).cpu().numpy().reshape(-1, 3)
    st = pt.stats()
    print(f"[transp] oracle, gpu stats {st}", flush=True)
    o = OraclePT(sc).render(sc["camera"], W, H, spp, 5, 21, threads=16)
    assert np.all(np.isfinite(g))
    close = np.abs(_lum(g) - _lum(o)) <= 1e-3 * np.maximum(_lum(o), 1e-3)
    e = float(np.mean((_lum(g) - _lum(o)) ** 2 / (_lum(o) ** 2 + 1e-3)))
    print(f"[transp] rel-MSE {e:.3e}, {close.mean() * 100:.2f}% within 1e-3", flush=True)
    assert e < 1e-4 and close.mean() > 0.97
    # the panels matter: the same scene with them opaque (white Lambert) renders differently
    opaque = dict(sc, materials=sc["materials"][:-1] + [I.lambert((0.7, 0.7, 0.7))])
    _, pt2 = _pt(gpu_ctx, opaque)
    g2 = pt2.render(I.camera(**sc["camera"]), W, H, spp, seed=21).cpu().numpy().reshape(-1, 3)
    assert np.mean(np.abs(_lum(g2) - _lum(g))) > 0.01
    # batched Li over caller rays follows the same paths
    rays = scenes.random_rays(4096, 5, lo=0.05, hi=0.95)
    keys = np.stack([np.arange(4096, dtype=np.uint32), np.full(4096, 2, np.uint32)], 1)
    gl = pt.li(rays, keys, seed=4).cpu().numpy()
    ol = OraclePT(sc).li(rays, keys, 5, 4, threads=16)
    ok = np.all(np.abs(gl - ol) <= 1e-3 * np.maximum(np.abs(ol), 1e-3), axis=1)
    assert ok.mean() > 0.99


def test_served_transparent_scene_in_the_mix(gpu_ctx):
    world = [panel_scene(), scenes.cornell_sphere()]
    built = []
    for sc in world:
        acc, pt = _pt(gpu_ctx, sc)
        built.append((acc, sc, pt))
    srv = S.Server(gpu_ctx, wave_paths=1 << 16, max_active=8)
    ids = [srv.add_scene(acc, sc["materials"], tri_material=sc["tri_material"], normals=sc["normals"],
                         quad_radiance=sc["quad_radiance"], env_radiance=sc["env"]) for acc, sc, _ in built]
    cam = I.camera(**world[0]["camera"])
    reqs = [(k % 2, 48 + 8 * (k % 3), 40, 3 + k, 5, 100 + k) for k in range(6)]
    print("[transp] served: 6 requests over a transparent and an opaque scene", flush=True)
    rids = [srv.submit(ids[k], cam, w, h, spp, depth, seed) for k, w, h, spp, depth, seed in reqs]
    for r in rids:
        assert srv.wait(r)["state"] == "done"
    for (k, w, h, spp, depth, seed), r in zip(reqs, rids):
        got = srv.result(r)
        ref = built[k][2].render(cam, w, h, spp, seed=seed, max_depth=depth).cpu().numpy().reshape(h, w, 3)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (k, w, h, spp)
    srv.close()
