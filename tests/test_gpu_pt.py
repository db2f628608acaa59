"""GPU path tracer vs the oracle's CPU path tracer.

With the same random-number keys the two follow the same paths, so images agree to float
rounding except for the rare path whose branch flips on a last-bit difference (libm vs device
transcendental functions): the bar is per-pixel relative MSE < 1e-4 and 97% of pixels within
1e-3 relative. With independent keys the images are compared as converged estimates
(relative MSE < 1e-3 against the oracle at 4096 spp)."""
import numpy as np
import pytest
import torch

from oracle.oracle import OraclePT
from sort_amd import Obvh, integrator as I, scenes

pytestmark = pytest.mark.gpu


def _gpu(ctx, sc, depth):
    acc = Obvh(ctx).build(sc["pos"], sc["idx"], sc["quads"])
    pt = I.PathTracing(acc, sc["materials"], tri_material=sc["tri_material"], normals=sc["normals"],
                       quad_radiance=sc["quad_radiance"], env_radiance=sc["env"], max_recursive_depth=depth)
    return acc, pt


def _lum(img):
    return 0.212671 * img[..., 0] + 0.715160 * img[..., 1] + 0.072169 * img[..., 2]


def rel_mse(img, ref, eps=1e-3):
    a, b = _lum(img).ravel(), _lum(ref).ravel()
    return float(np.mean((a - b) ** 2 / (b * b + eps)))


def _render_both(ctx, sc, W, H, spp, depth, seed, crop_cam=None):
    cam = crop_cam or sc["camera"]
    acc, pt = _gpu(ctx, sc, depth)
    print(f"[pt] gpu render {W}x{H} spp={spp} depth={depth}", flush=True)
    g = pt.render(I.camera(**cam), W, H, spp, seed=seed).cpu().numpy().reshape(-1, 3)
    st = pt.stats()
    print(f"[pt] oracle render, gpu stats {st}", flush=True)
    o = OraclePT(sc).render(cam, W, H, spp, depth, seed, threads=16)
    return g, o, st


SPHERE_MATS = {
    "lambert": None,
    "ggx": I.ggx((0.95, 0.9, 0.8), 0.3),
    "dielectric": I.dielectric((1, 1, 1), (1, 1, 1), roughness=0.15, ior=1.5),
    "disney": I.disney((0.8, 0.4, 0.2), metallic=0.4, specular=0.5, roughness=0.35, sheen=0.5, clearcoat=0.6,
                       clearcoat_gloss=0.8),
}


@pytest.mark.parametrize("mat", list(SPHERE_MATS))
def test_cornell_same_keys_matches_oracle(gpu_ctx, mat):
    sc = scenes.cornell_sphere(sphere_material=SPHERE_MATS[mat])
    g, o, st = _render_both(gpu_ctx, sc, 64, 64, 8, 5, seed=7)
    assert np.all(np.isfinite(g))
    assert st["extension_rays"] >= 64 * 64 * 8 and st["shadow_rays"] > 0
    close = np.abs(_lum(g) - _lum(o)) <= 1e-3 * np.maximum(_lum(o), 1e-3)
    print(f"[pt] {mat}: rel-MSE {rel_mse(g, o):.3e}, {close.mean() * 100:.2f}% pixels within 1e-3", flush=True)
    assert rel_mse(g, o) < 1e-4
    assert close.mean() > 0.97


def test_cornell_full_config_matches_oracle(gpu_ctx):
    """Config 1 at full size: 256^2, 16 spp, depth 5, NEE."""
    sc = scenes.cornell_sphere()
    g, o, st = _render_both(gpu_ctx, sc, 256, 256, 16, 5, seed=1)
    print(f"[pt] C1 rel-MSE {rel_mse(g, o):.3e}", flush=True)
    assert rel_mse(g, o) < 1e-4


def test_spheres_scene_env_and_quads_matches_oracle(gpu_ctx):
    """Config 3's material mix (Lambert / GGX / dielectric), environment + 4 quad lights, depth 8
    with Russian roulette, on a reduced scene (40 icospheres of level 3)."""
    sc = scenes.spheres_scene(n_spheres=40, level=3)
    g, o, _ = _render_both(gpu_ctx, sc, 96, 64, 4, 8, seed=3)
    close = np.abs(_lum(g) - _lum(o)) <= 1e-3 * np.maximum(_lum(o), 1e-3)
    print(f"[pt] spheres rel-MSE {rel_mse(g, o):.3e}, {close.mean() * 100:.2f}% within 1e-3", flush=True)
    assert rel_mse(g, o) < 1e-4
    assert close.mean() > 0.97


def test_li_matches_oracle_per_ray(gpu_ctx):
    sc = scenes.cornell_sphere(sphere_material=SPHERE_MATS["disney"])
    acc, pt = _gpu(gpu_ctx, sc, 6)
    rays = scenes.random_rays(20000, 77, lo=0.05, hi=0.95)
    keys = np.stack([np.arange(20000, dtype=np.uint32), np.full(20000, 3, np.uint32)], 1)
    g = pt.li(rays, keys, seed=5).cpu().numpy()
    o = OraclePT(sc).li(rays, keys, 6, 5, threads=16)
    close = np.all(np.abs(g - o) <= 1e-3 * np.maximum(np.abs(o), 1e-3), axis=1)
    print(f"[pt] li: {close.mean() * 100:.2f}% rays within 1e-3", flush=True)
    assert close.mean() > 0.99


def test_converged_image_independent_keys(gpu_ctx):
    """GPU at 16384 spp vs the oracle at 4096 spp with independent keys: relative MSE < 1e-3."""
    sc = scenes.cornell_sphere()
    W = H = 24
    acc, pt = _gpu(gpu_ctx, sc, 5)
    g = pt.render(I.camera(**sc["camera"]), W, H, 16384, seed=101).cpu().numpy().reshape(-1, 3)
    o = OraclePT(sc).render(sc["camera"], W, H, 4096, 5, 202, threads=16)
    e = rel_mse(g, o)
    print(f"[pt] converged rel-MSE {e:.3e}", flush=True)
    assert e < 1e-3


def test_render_is_deterministic_and_batch_independent(gpu_ctx):
    sc = scenes.cornell_sphere(sphere_material=SPHERE_MATS["dielectric"])
    acc, pt = _gpu(gpu_ctx, sc, 5)
    cam = I.camera(**sc["camera"])
    a = pt.render(cam, 48, 40, 12, seed=9).cpu().numpy()
    b = pt.render(cam, 48, 40, 12, seed=9).cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    c = pt.render(cam, 48, 40, 12, seed=10).cpu().numpy()
    assert not np.array_equal(a, c)


def test_c1_ray_dump_first_million_rays(gpu_ctx):
    """Config 1's ray dump: the first 1,048,576 camera + bounce rays of the 256^2, 16 spp, depth 5
    Cornell render in (pixel, sample, bounce) order with their closest hits. Every GPU hit equals,
    bit for bit, the oracle BVH's hit for the same ray; camera rays equal the oracle path tracer's
    bit for bit; later rays equal the oracle's wherever the two paths still coincide (a path may
    part from the oracle's after a last-bit difference of a transcendental function)."""
    import torch
    sc = scenes.cornell_sphere()
    W = H = 256
    spp, depth, seed = 16, 5, 1
    pix_end = 64 * W  # 1,310,720 slots, more than 1M of them traced
    acc, pt = _gpu(gpu_ctx, sc, depth)
    cam = I.camera(**sc["camera"])
    print("[pt] gpu ray dump", flush=True)
    g_rays, g_hits, g_valid = pt.ray_dump(cam, W, H, spp, 0, pix_end, seed=seed, max_depth=depth)
    g_rays, g_hits, g_valid = g_rays.cpu().numpy(), g_hits.cpu().numpy(), g_valid.cpu().numpy().astype(bool)
    print("[pt] oracle ray dump", flush=True)
    o = OraclePT(sc)
    o_rays, o_hits, o_valid = o.ray_dump(sc["camera"], W, H, spp, depth, seed, 0, pix_end, threads=16)
    # the first 1,048,576 traced rays in slot order
    idx = np.flatnonzero(g_valid)[: 1 << 20]
    assert len(idx) == 1 << 20
    again = o.bvh.closest(g_rays[idx], threads=16)
    assert np.array_equal(again.view(np.uint32), g_hits[idx].view(np.uint32))
    cam0 = (np.arange(len(g_valid)) % depth) == 0
    assert np.array_equal(g_valid[cam0], o_valid[cam0]) and g_valid[cam0].all()
    assert np.array_equal(g_rays[cam0].view(np.uint32), o_rays[cam0].view(np.uint32))
    assert np.array_equal(g_hits[cam0].view(np.uint32), o_hits[cam0].view(np.uint32))
    both = g_valid & o_valid
    same = both & np.all(g_rays.view(np.uint32) == o_rays.view(np.uint32), axis=1)
    err = np.abs(g_rays[both] - o_rays[both]).max(axis=1)
    close = err <= 1e-4
    prim_eq = g_hits[both][:, 1].view(np.uint32) == o_hits[both][:, 1].view(np.uint32)
    print(f"[pt] traced slots gpu {g_valid.sum()} oracle {o_valid.sum()}, both {both.sum()}; rays bitwise equal "
          f"{same.sum() / both.sum() * 100:.3f}%, within 1e-4 {close.mean() * 100:.4f}% (max {err.max():.3g}); "
          f"same hit primitive {prim_eq.mean() * 100:.4f}%", flush=True)
    assert both.sum() >= 0.999 * o_valid.sum()
    assert close.mean() >= 0.999
    assert prim_eq.mean() >= 0.995
    # bit-identical rays have bit-identical hits (the traversal is exact)
    assert np.array_equal(g_hits[same].view(np.uint32), o_hits[same].view(np.uint32))


def test_per_sample_clamp_matches_oracle(gpu_ctx):
    """The scene's clamp (ImageEvaluation's clamping) on the GPU: the same keys give the oracle's
    clamped image; the light seen head-on is exactly the clamp; clamp 0 is the unclamped render;
    a negative clamp is refused. A scene file's clamping reaches the render through the loader."""
    from oracle import scene_file_oracle as SF
    from sort_amd import SortGpuError, scene_file
    sc = dict(scenes.cornell_sphere(), clamp=0.5)
    acc = Obvh(gpu_ctx).build(sc["pos"], sc["idx"], sc["quads"])

    def pt(clamp):
        return I.PathTracing(acc, sc["materials"], tri_material=sc["tri_material"], normals=sc["normals"],
                             quad_radiance=sc["quad_radiance"], env_radiance=sc["env"], max_recursive_depth=5,
                             clamp=clamp)
    print("[pt] clamp 0.5", flush=True)
    g = pt(0.5).render(I.camera(**sc["camera"]), 64, 64, 8, seed=7).cpu().numpy().reshape(-1, 3)
    o = OraclePT(sc).render(sc["camera"], 64, 64, 8, 5, 7, threads=16)
    assert np.all(g <= np.float32(0.5)) and np.all(g >= 0)
    close = np.abs(_lum(g) - _lum(o)) <= 1e-3 * np.maximum(_lum(o), 1e-3)
    assert close.mean() > 0.97 and rel_mse(g, o) < 1e-4
    up = dict(eye=(0.5, 0.9, 0.5), target=(0.5, 1.0, 0.5), up=(0.0, 0.0, 1.0), fov_y_deg=5.0)
    head_on = pt(5.0).render(I.camera(**up), 16, 16, 4, seed=3).cpu().numpy()
    assert np.all(head_on == np.float32(5.0))
    a = pt(0.0).render(I.camera(**sc["camera"]), 32, 32, 4, seed=9).cpu().numpy()
    b = I.PathTracing(acc, sc["materials"], tri_material=sc["tri_material"], normals=sc["normals"],
                      quad_radiance=sc["quad_radiance"], env_radiance=sc["env"], max_recursive_depth=5).render(
        I.camera(**sc["camera"]), 32, 32, 4, seed=9).cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    with pytest.raises(SortGpuError, match="clamp"):
        pt(-1.0)
    # the loader's clamping reaches the render
    data = bytearray(SF.export_scene_dict(scenes.cornell_sphere(), width=32, height=32, spp=4))
    # clamping is the float after version, resource path, threads, spp, width and height
    k = 4 + len(b"/tmp/sort/\0") + 16
    data[k:k + 4] = np.float32(0.5).tobytes()
    loaded = scene_file.parse(bytes(data))
    assert loaded["clamp"] == np.float32(0.5)
    g = I.PathTracing(acc, loaded["materials"], tri_material=loaded["tri_material"], normals=loaded["normals"],
                      quad_radiance=loaded["quad_radiance"], env_radiance=loaded["env"], max_recursive_depth=5,
                      clamp=loaded["clamp"]).render(I.camera(**loaded["camera"]), 32, 32, 4, seed=5).cpu().numpy()
    assert np.all(g <= np.float32(0.5))


@pytest.mark.parametrize("mat", ["lambert", "dielectric"])
def test_shade_order_changes_no_result(gpu_ctx, mat):
    """Shading each bounce's hits in spatial order (the bucket sort before k_shade) changes only
    speed: the image is bitwise the one shaded in queue order."""
    from sort_amd.batch import set_tuning
    sc = scenes.spheres_scene(n_spheres=40, level=3) if mat == "lambert" else \
        scenes.cornell_sphere(sphere_material=SPHERE_MATS["dielectric"])
    acc, pt = _gpu(gpu_ctx, sc, 8)
    cam = I.camera(**sc["camera"])
    imgs = []
    try:
        for order in (1, 2, 0):  # never, always, default (a 128x128x32 wave is 2^19 paths: sorted)
            set_tuning(gpu_ctx, shade_order=order)
            print(f"[pt] shade_order={order}", flush=True)
            imgs.append(pt.render(cam, 128, 128, 32, seed=5).cpu().numpy())
    finally:
        set_tuning(gpu_ctx)
    assert np.array_equal(imgs[0].view(np.uint32), imgs[1].view(np.uint32))
    assert np.array_equal(imgs[0].view(np.uint32), imgs[2].view(np.uint32))


def test_stage_timing_changes_no_result(gpu_ctx):
    """Stage timing (events around each bounce's traversal, shading and shadow rays) leaves the
    image bitwise unchanged and accounts for every bounce of the frame's waves."""
    sc = scenes.cornell_sphere()
    acc, pt = _gpu(gpu_ctx, sc, 5)
    cam = I.camera(**sc["camera"])
    a = pt.render(cam, 64, 64, 8, seed=3).cpu().numpy()
    assert pt.stage_times()["bounces"] == 0
    pt.set_timing(True)
    b = pt.render(cam, 64, 64, 8, seed=3).cpu().numpy()
    t = pt.stage_times()
    print(f"[pt] stage times {t}", flush=True)
    pt.set_timing(False)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert t["bounces"] >= 5 and t["bounces"] % 5 == 0
    assert t["extend_ms"] > 0 and t["shade_ms"] > 0 and t["shadow_ms"] > 0
    total_ms = pt.stats()["gpu_us"] / 1e3
    assert t["extend_ms"] + t["shade_ms"] + t["shadow_ms"] <= 2 * total_ms + 1.0
