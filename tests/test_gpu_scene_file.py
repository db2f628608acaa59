"""A scene read from a .sort file renders on the GPU like the oracle renders the same loaded scene
(its smoothed tangents included), and, without the tangents, like the built-in scene it was
written from (same geometry and lights; normals renormalised).
"""
import numpy as np
import pytest

from oracle import scene_file_oracle as O
from oracle.oracle import OraclePT
from sort_amd import Obvh, integrator as I, scene_file, scenes

pytestmark = pytest.mark.gpu


def _lum(img):
    return 0.212671 * img[..., 0] + 0.715160 * img[..., 1] + 0.072169 * img[..., 2]


def _render(ctx, sc, W, H, spp, depth, seed):
    acc = Obvh(ctx).build(sc["pos"], sc["idx"], sc["quads"])
    pt = I.PathTracing(acc, sc["materials"], tri_material=sc["tri_material"], normals=sc["normals"],
                       tangents=sc.get("tangents"), quad_radiance=sc["quad_radiance"], env_radiance=sc["env"],
                       max_recursive_depth=depth)
    img = pt.render(I.camera(**sc["camera"]), W, H, spp, seed=seed).cpu().numpy().reshape(-1, 3)
    pt.close()
    return img


@pytest.mark.parametrize("sphere", ["lambert", "ggx"])
def test_loaded_scene_renders_like_oracle_and_source(gpu_ctx, tmp_path, sphere):
    mat = None if sphere == "lambert" else I.ggx((0.95, 0.9, 0.8), 0.3)
    src = scenes.cornell_sphere(sphere_material=mat)
    path = tmp_path / "cornell.sort"
    path.write_bytes(O.export_scene_dict(src, width=64, height=64, spp=8, max_depth=5))
    sc = scene_file.load(path)
    cfg = sc["config"]
    W, H, spp, depth = cfg["width"], cfg["height"], cfg["spp"], cfg["max_recursive_depth"]
    print(f"[scene file] {sphere}: {sc['ntris']} tris, {len(sc['quads'])} quads, {W}x{H} spp={spp}", flush=True)
    g = _render(gpu_ctx, sc, W, H, spp, depth, seed=11)
    o = OraclePT(sc).render(sc["camera"], W, H, spp, depth, 11, threads=16)
    assert np.all(np.isfinite(g))
    close = np.abs(_lum(g) - _lum(o)) <= 1e-3 * np.maximum(_lum(o), 1e-3)
    assert close.mean() > 0.97
    a, b = _lum(g), _lum(o)
    assert float(np.mean((a - b) ** 2 / (b * b + 1e-3))) < 1e-4
    # the scene it was written from (edge tangents): same keys, so the same paths up to the
    # renormalised normals when the loaded scene drops its tangents too
    s = _render(gpu_ctx, src, W, H, spp, depth, seed=11)
    e = _render(gpu_ctx, {k: v for k, v in sc.items() if k != "tangents"}, W, H, spp, depth, seed=11)
    a, b = _lum(e), _lum(s)
    assert float(np.mean((a - b) ** 2 / (b * b + 1e-3))) < 1e-3
