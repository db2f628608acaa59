"""The benchmark's workloads at their full sizes, checked for correctness.

* 2M-triangle path tracer: the spheres scene (400 icospheres of level 5 on a ground plane, 50%
  Lambert / 30% GGX / 20% dielectric, environment + 4 quad lights) at 1024 x 1024, 64 spp, depth 8
  with Russian roulette, exactly as timed; rows 480..543 rendered by the oracle with the same
  random-number keys must agree to per-pixel relative MSE < 1e-4 with 97% of pixels within 1e-3.
* Cornell box at 256 x 256 converged: the GPU at 16384 spp against the oracle at 4096 spp with
  independent keys, relative MSE < 1e-3.
* Batched form: the 64-scene workload (16.0M triangles, 25.1M incoherent rays) in one pooled launch
  equals one launch per scene bit for bit on every ray, and the oracle on every ray of the smallest,
  the median and the largest scene.
* Served form: the 16 served scenes under a seeded Poisson trace of 20 requests (20 per second),
  with an HBM budget that holds only part of the scenes so least-recently-used scenes are evicted
  and reloaded; every served image is bitwise the single-request render of the same request.
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleBvh, OraclePT
from sort_amd import Obvh, batch, integrator as I, scenes
from sort_amd import serve as S

pytestmark = pytest.mark.gpu


def _lum(img):
    return 0.212671 * img[..., 0] + 0.715160 * img[..., 1] + 0.072169 * img[..., 2]


def _rel_mse(img, ref, eps=1e-3):
    a, b = _lum(img).ravel(), _lum(ref).ravel()
    return float(np.mean((a - b) ** 2 / (b * b + eps)))


def _pt(ctx, sc, depth):
    acc = Obvh(ctx).build(sc["pos"], sc["idx"], sc["quads"])
    pt = I.PathTracing(acc, sc["materials"], tri_material=sc.get("tri_material"), normals=sc.get("normals"),
                       quad_radiance=sc["quad_radiance"], env_radiance=sc["env"], max_recursive_depth=depth)
    return acc, pt


def test_spheres_2m_full_frame_crop_same_keys(gpu_ctx):
    sc = scenes.spheres_scene()
    assert len(sc["idx"]) == 2048002
    W = H = 1024
    spp, depth, seed, row0, rows = 64, 8, 1, 480, 64
    acc, pt = _pt(gpu_ctx, sc, depth)
    print(f"[full] spheres: gpu {W}x{H} {spp} spp depth {depth}", flush=True)
    g = pt.render(I.camera(**sc["camera"]), W, H, spp, seed=seed, max_depth=depth).cpu().numpy().reshape(-1, 3)
    st = pt.stats()
    assert st["paths"] == W * H * spp
    print(f"[full] spheres: oracle rows {row0}..{row0 + rows - 1}, gpu stats {st}", flush=True)
    o = OraclePT(sc).render(sc["camera"], W, H, spp, depth, seed, threads=16, crop=(W * row0, W * (row0 + rows)))
    crop = g[W * row0: W * (row0 + rows)]
    assert np.all(np.isfinite(crop))
    close = np.abs(_lum(crop) - _lum(o)) <= 1e-3 * np.maximum(_lum(o), 1e-3)
    e = _rel_mse(crop, o)
    print(f"[full] spheres crop: rel-MSE {e:.3e}, {close.mean() * 100:.2f}% pixels within 1e-3", flush=True)
    assert e < 1e-4
    assert close.mean() > 0.97
    pt.close()


def test_cornell_256_converged_independent_keys(gpu_ctx):
    sc = scenes.cornell_sphere()
    W = H = 256
    acc, pt = _pt(gpu_ctx, sc, 5)
    print("[full] cornell 256^2: gpu 16384 spp", flush=True)
    g = pt.render(I.camera(**sc["camera"]), W, H, 16384, seed=101, max_depth=5).cpu().numpy().reshape(-1, 3)
    print("[full] cornell 256^2: oracle 4096 spp", flush=True)
    o = OraclePT(sc).render(sc["camera"], W, H, 4096, 5, 202, threads=16)
    e = _rel_mse(g, o)
    print(f"[full] cornell 256^2 converged rel-MSE {e:.3e}", flush=True)
    assert e < 1e-3
    pt.close()


def test_batched_64_scenes_full_workload(gpu_ctx):
    print("[full] batched: building 64 scenes", flush=True)
    w = batch.build_workload(gpu_ctx, n_scenes=64, width=8)
    n = len(w["rays"])
    assert len(w["accels"]) == 64 and n > 20_000_000
    pool = batch.Pool(gpu_ctx, w["accels"])
    d_rays = torch.from_numpy(w["rays"]).cuda()
    d_sid = torch.from_numpy(w["scene_of_ray"].view(np.int32)).cuda()
    print(f"[full] batched: pooled launch over {n} rays", flush=True)
    got = pool.trace(d_rays, d_sid).cpu().numpy()
    pool.check_errors()
    off = w["offsets"]
    for i, acc in enumerate(w["accels"]):
        ref = acc.get_intersect(d_rays[off[i]:off[i + 1]]).cpu().numpy()
        assert np.array_equal(got[off[i]:off[i + 1]].view(np.uint32), ref.view(np.uint32)), f"scene {i}"
    counts = np.diff(off)
    order = np.argsort(w["tris"], kind="stable")
    for i in {int(order[0]), int(order[len(order) // 2]), int(order[-1])}:
        pos, idx = w["geometry"][i]
        print(f"[full] batched: oracle on scene {i} ({w['tris'][i]} tris, {counts[i]} rays)", flush=True)
        ref = OracleBvh(pos, idx).closest(w["rays"][off[i]:off[i + 1]], threads=16)
        assert np.array_equal(got[off[i]:off[i + 1]].view(np.uint32), ref.view(np.uint32)), f"oracle scene {i}"
    pool.close()


def test_served_poisson_trace_with_evictions(gpu_ctx):
    built = S.build_scenes(gpu_ctx)
    trace = S.request_trace(20)
    # scene sizes on the device, measured one at a time
    probe = S.Server(gpu_ctx, wave_paths=1 << 23, manual=True)  # the served run's workspace
    ids = S.add_scenes(probe, built)
    fixed = probe.stats()["workspace_bytes"]
    sizes = []
    for i in ids:
        before = probe.stats()["bytes_in_use"]
        probe.wait(probe.submit(i, S.random_camera(np.random.default_rng(i)), 8, 8, 1))
        sizes.append(probe.stats()["bytes_in_use"] - before)
    probe.close()
    sizes = np.array(sizes)
    films = 4 * 1024 * 1024 * 16
    # room for half of the scenes the trace uses (at least its largest): every request fits
    # alone, but its scenes do not fit together, so least-recently-used ones must be evicted
    used = sizes[sorted({r["scene"] for r in trace})]
    budget = int(fixed + films + max(used.max(), used.sum() // 2))
    assert len(used) >= 4 and budget < fixed + films + used.sum()
    srv = S.Server(gpu_ctx, hbm_budget=budget, wave_paths=1 << 23, max_active=4)
    ids = S.add_scenes(srv, built)
    print(f"[full] served: 20 Poisson requests, budget {budget} ({len(built)} scenes, {sizes.sum()} scene bytes)",
          flush=True)
    rids, out, wall = S.run_served(srv, trace, ids)
    st = srv.stats()
    print(f"[full] served: wall {wall:.2f}s, stats {st}", flush=True)
    assert all(o["state"] == "done" for o in out), [o["state"] for o in out]
    assert st["evictions"] > 0 and st["peak_bytes"] <= budget
    served = [srv.result(r) for r in rids]
    srv.close()
    pts = {}
    for k, (r, img) in enumerate(zip(trace, served)):
        if r["scene"] not in pts:
            acc, sc = built[r["scene"]]
            pts[r["scene"]] = I.PathTracing(acc, sc["materials"], tri_material=sc["tri_material"],
                                            quad_radiance=sc["quad_radiance"], env_radiance=sc["env"])
        ref = pts[r["scene"]].render(r["cam"], r["width"], r["height"], r["spp"], seed=r["seed"],
                                     max_depth=r["depth"]).cpu().numpy().reshape(r["height"], r["width"], 3)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), f"request {k}: {r['width']}^2 {r['spp']} spp"
    for pt in pts.values():
        pt.close()
