"""Served form on the GPU: every request's image equals, bit for bit, the single-request render of
the same scene, camera, size, spp, depth and seed -- alone, in a mix of concurrent requests over
several resident scenes, under LRU eviction, and next to cancelled requests. The single-request
render is itself checked against the oracle's CPU path tracer (test_gpu_pt.py); one served request
is checked against the oracle directly here."""
import numpy as np
import pytest

from oracle.oracle import OraclePT
from sort_amd import Obvh, integrator as I, scenes
from sort_amd import serve as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(gpu_ctx):
    """Cornell box + 3 small soups: (accel, scene dict, PathTracing) each."""
    out = []
    sc = scenes.cornell_sphere()
    out.append(sc)
    for i, n in enumerate([20_000, 60_000, 5_000]):
        out.append(S.soup_scene(n, seed=4242 + i))
    res = []
    for sc in out:
        acc = Obvh(gpu_ctx).build(sc["pos"], sc["idx"], sc["quads"])
        pt = I.PathTracing(acc, sc["materials"], tri_material=sc["tri_material"], normals=sc.get("normals"),
                           quad_radiance=sc["quad_radiance"], env_radiance=sc["env"])
        res.append((acc, sc, pt))
    return res


def _server(gpu_ctx, world, **kw):
    srv = S.Server(gpu_ctx, **kw)
    ids = [srv.add_scene(acc, sc["materials"], tri_material=sc["tri_material"], normals=sc.get("normals"),
                         quad_radiance=sc["quad_radiance"], env_radiance=sc["env"]) for acc, sc, _ in world]
    return srv, ids


def _cam(world, k, g):
    if k == 0:
        return I.camera(**world[0][1]["camera"])
    return S.random_camera(g)


def _plain(world, k, cam, w, h, spp, depth, seed):
    return world[k][2].render(cam, w, h, spp, seed=seed, max_depth=depth).cpu().numpy()


def _mix(world, n, seed):
    g = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        k = int(g.integers(0, len(world)))
        w = int(g.choice([16, 48, 64, 100]))
        h = int(g.choice([16, 40, 64]))
        reqs.append(dict(k=k, cam=_cam(world, k, g), w=w, h=h, spp=int(g.choice([1, 3, 8, 16])),
                         depth=int(g.choice([1, 5, 8])), seed=int(g.integers(1, 1 << 40))))
    return reqs


def test_served_request_matches_plain_render_and_oracle(gpu_ctx, world):
    srv, ids = _server(gpu_ctx, world)
    cam = I.camera(**world[0][1]["camera"])
    print("[serve] cornell 64x64 8spp depth 5", flush=True)
    rid = srv.submit(ids[0], cam, 64, 64, 8, max_depth=5, seed=7)
    info = srv.wait(rid, 60.0)
    assert info["state"] == "done", (info, srv.message(rid))
    got = srv.result(rid)
    ref = _plain(world, 0, cam, 64, 64, 8, 5, 7)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    o = OraclePT(world[0][1]).render(world[0][1]["camera"], 64, 64, 8, 5, 7, threads=16).reshape(64, 64, 3)
    lum = lambda x: 0.212671 * x[..., 0] + 0.715160 * x[..., 1] + 0.072169 * x[..., 2]
    a, b = lum(got).ravel(), lum(o).ravel()
    assert float(np.mean((a - b) ** 2 / (b * b + 1e-3))) < 1e-4
    st = srv.stats()
    assert st["done"] == 1 and st["extension_rays"] >= 64 * 64 * 8
    srv.close()


@pytest.mark.parametrize("wave_paths,graphs,streams", [(1 << 14, False, 2), (1 << 18, False, 2), (1 << 14, True, 2),
                                                       (1 << 16, True, 3), (1 << 16, False, 1)])
def test_alone_equals_in_mix(gpu_ctx, world, wave_paths, graphs, streams):
    """Graph mode replays each captured bounce loop for many waves, with the server's own eager
    work (descriptor copies, film clears and updates, result copies) between the replays; every
    result is checked against the eager single-request render."""
    reqs = _mix(world, 16, seed=wave_paths + streams)
    srv, ids = _server(gpu_ctx, world, wave_paths=wave_paths, max_active=8, graphs=graphs, inflight_waves=streams)
    print(f"[serve] mix of {len(reqs)} requests, wave {wave_paths}, graphs {graphs}, streams {streams}", flush=True)
    rids = [srv.submit(ids[r["k"]], r["cam"], r["w"], r["h"], r["spp"], r["depth"], r["seed"]) for r in reqs]
    for rid in rids:
        info = srv.wait(rid, 120.0)
        assert info["state"] == "done", (info, srv.message(rid))
    st = srv.stats()
    print(f"[serve] stats {st}", flush=True)
    assert st["waves"] < sum(r["spp"] for r in reqs)  # requests shared waves
    for r, rid in zip(reqs, rids):
        ref = _plain(world, r["k"], r["cam"], r["w"], r["h"], r["spp"], r["depth"], r["seed"])
        assert np.array_equal(srv.result(rid).view(np.uint32), ref.view(np.uint32)), r
    srv.close()


def test_eviction_and_cancellation_leave_results_unchanged(gpu_ctx, world):
    probe, ids = _server(gpu_ctx, world, wave_paths=1 << 14, manual=True)
    fixed = probe.stats()["workspace_bytes"]
    sizes = []
    for i in ids:
        before = probe.stats()["bytes_in_use"]
        probe.submit(i, S.random_camera(np.random.default_rng(0)), 8, 8, 1)
        probe.drain()
        sizes.append(probe.stats()["bytes_in_use"] - before)
    probe.close()
    budget = fixed + sum(sorted(sizes)[-2:]) + 2 * 100 * 64 * 16
    reqs = _mix(world, 12, seed=99)
    srv, ids = _server(gpu_ctx, world, wave_paths=1 << 14, max_active=2, hbm_budget=budget)
    print(f"[serve] budget {budget} scene bytes {sizes}", flush=True)
    g = np.random.default_rng(5)
    long_cam = S.random_camera(g)
    victim = srv.submit(ids[2], long_cam, 64, 64, 512, 8, 1234)  # long; cancelled mid-run
    rids = [srv.submit(ids[r["k"]], r["cam"], r["w"], r["h"], r["spp"], r["depth"], r["seed"]) for r in reqs]
    srv.cancel(victim)
    for rid in rids:
        info = srv.wait(rid, 120.0)
        assert info["state"] == "done", (info, srv.message(rid))
    assert srv.info(victim)["state"] == "cancelled"
    st = srv.stats()
    print(f"[serve] stats {st}", flush=True)
    assert st["evictions"] > 0 and st["peak_bytes"] <= budget
    for r, rid in zip(reqs, rids):
        ref = _plain(world, r["k"], r["cam"], r["w"], r["h"], r["spp"], r["depth"], r["seed"])
        assert np.array_equal(srv.result(rid).view(np.uint32), ref.view(np.uint32)), r
    srv.close()


def test_multi_wave_render_equals_served(gpu_ctx, world):
    """A frame the single-request render splits into waves (1024^2 at 8 spp: three waves of 3, 3 and
    2 samples on its three streams) equals, bit for bit, the server's result for the same request
    (one 8-sample wave)."""
    cam = I.camera(**world[0][1]["camera"])
    plain = _plain(world, 0, cam, 1024, 1024, 8, 5, 77)
    srv, ids = _server(gpu_ctx, world)
    rid = srv.submit(ids[0], cam, 1024, 1024, 8, max_depth=5, seed=77)
    assert srv.wait(rid, 120.0)["state"] == "done"
    assert np.array_equal(srv.result(rid).view(np.uint32), plain.view(np.uint32))
    st = world[0][2].stats()
    assert st["paths"] == 1024 * 1024 * 8 and st["extension_rays"] > st["paths"]
    srv.close()
