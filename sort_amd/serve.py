"""Served form: a continuous-batching render server (``sortgpu_server_*``) and the workload it is
measured on.

The server keeps scenes, the wave workspace and request films resident in HBM, admits requests
in arrival order under an HBM budget (evicting least-recently-used idle scenes), and packs whole
samples of many requests into each wave of the bounce loop. A request's image is bitwise what
``PathTracing.render`` gives for the same scene, camera, size, spp, depth and seed.

Workload (seeded with ``scenes.SEED_SERVE``): 16 scenes, seeded triangle soups like the 1M-triangle
soup with counts ~ LogNormal(ln 150K, 1.0) clipped to [50K, 2M], materials 50% Lambert / 30% GGX /
20% dielectric per triangle, a constant environment and one quad light; requests arrive as a
Poisson process (20 per second), each = uniform scene + random camera + resolution from
{256^2, 512^2, 1024^2} + spp from {4, 16, 64}, depth 8.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib, scenes
from ._lib import check, lib
from .integrator import camera, dielectric, ggx, lambert, scene_desc

STATES = {0: "queued", 1: "active", 2: "done", 3: "failed", 4: "cancelled", 5: "timed_out"}
STAT_KEYS = ["waves", "paths", "extension_rays", "shadow_rays", "scene_loads", "evictions", "resident_scenes",
             "bytes_in_use", "peak_bytes", "done", "failed", "cancelled", "timed_out", "device_busy_us", "budget",
             "workspace_bytes"]


def _info(i: _lib.RequestInfo) -> dict:
    return {"state": STATES.get(i.state, str(i.state)), "status": i.status, "samples_done": i.samples_done,
            "spp": i.spp, "t_submit": i.t_submit, "t_admit": i.t_admit, "t_done": i.t_done}


class Server:
    """Render server on a device context (or, with host_device, a host-only stand-in device)."""

    def __init__(self, ctx, hbm_budget=0, wave_paths=0, max_active=0, inflight_waves=0, manual=False,
                 host_device=False, fail_alloc_at=0, graphs=False):
        self.ctx = ctx
        o = _lib.ServerOpts(int(hbm_budget), int(wave_paths), int(max_active), int(inflight_waves), int(manual),
                            int(host_device), int(fail_alloc_at), int(graphs))
        h = ctypes.c_void_p()
        check(lib().sortgpu_server_create(ctx.handle, ctypes.byref(o), ctypes.byref(h)), ctx.handle)
        self.handle = h
        self._sizes = {}

    def _chk(self, rc):
        return check(rc, self.ctx.handle)

    def add_scene(self, accel, materials, tri_material=None, normals=None, tangents=None, quad_radiance=None,
                  env_radiance=(0.0, 0.0, 0.0), clamp=0.0) -> int:
        d, keep = scene_desc(materials, tri_material, normals, tangents, quad_radiance, env_radiance, clamp)
        sid = ctypes.c_uint32()
        self._chk(lib().sortgpu_server_add_scene(self.handle, accel.handle, ctypes.byref(d), ctypes.byref(sid)))
        del keep
        return int(sid.value)

    def submit(self, scene, cam, width, height, spp, max_depth=8, seed=1, timeout_s=0.0) -> int:
        r = _lib.Request()
        r.scene, r.width, r.height, r.spp, r.max_depth = int(scene), int(width), int(height), int(spp), int(max_depth)
        r.seed = int(seed)
        r.camera = cam
        r.timeout_s = float(timeout_s)
        rid = ctypes.c_uint64()
        self._chk(lib().sortgpu_server_submit(self.handle, ctypes.byref(r), ctypes.byref(rid)))
        self._sizes[int(rid.value)] = (int(height), int(width))
        return int(rid.value)

    def cancel(self, rid):
        self._chk(lib().sortgpu_server_cancel(self.handle, rid))

    def info(self, rid) -> dict:
        i = _lib.RequestInfo()
        self._chk(lib().sortgpu_server_info(self.handle, rid, ctypes.byref(i)))
        return _info(i)

    def wait(self, rid, timeout_s=-1.0) -> dict:
        i = _lib.RequestInfo()
        self._chk(lib().sortgpu_server_wait(self.handle, rid, float(timeout_s), ctypes.byref(i)))
        return _info(i)

    def result(self, rid) -> np.ndarray:
        h, w = self._sizes[rid]
        out = np.empty((h, w, 3), np.float32)
        self._chk(lib().sortgpu_server_result(self.handle, rid, out.ctypes.data, out.size))
        return out

    def message(self, rid) -> str:
        buf = ctypes.create_string_buffer(512)
        self._chk(lib().sortgpu_server_message(self.handle, rid, buf, 512))
        return buf.value.decode()

    def release(self, rid):
        self._chk(lib().sortgpu_server_release(self.handle, rid))
        self._sizes.pop(rid, None)

    def step(self) -> bool:
        busy = ctypes.c_int()
        self._chk(lib().sortgpu_server_step(self.handle, ctypes.byref(busy)))
        return bool(busy.value)

    def drain(self, timeout_s=-1.0):
        self._chk(lib().sortgpu_server_drain(self.handle, float(timeout_s)))

    def stats(self) -> dict:
        out = np.zeros(16, np.uint64)
        self._chk(lib().sortgpu_server_stats(self.handle, out.ctypes.data))
        return dict(zip(STAT_KEYS, [int(x) for x in out]))

    def close(self):
        if self.handle:
            lib().sortgpu_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- workload ------------------------------------------------------------------------------------

RESOLUTIONS = (256, 512, 1024)
SPPS = (4, 16, 64)


def scene_sizes(n_scenes=16, seed=scenes.SEED_SERVE):
    g = scenes.rng(seed)
    return np.clip(np.exp(g.normal(np.log(150_000), 1.0, n_scenes)), 50_000, 2_000_000).astype(np.int64)


def soup_scene(n_tris, seed):
    """A seeded soup in [0,1]^3 with per-triangle materials (50/30/20 Lambert / GGX / dielectric
    from a palette of 8), a constant environment and one 0.5 x 0.5 quad light above it."""
    pos, idx = scenes.soup(int(n_tris), seed=seed)
    g = scenes.rng(seed + 77)
    materials = [lambert((0.7, 0.7, 0.7)), lambert((0.8, 0.3, 0.2)), lambert((0.2, 0.6, 0.3)), lambert((0.3, 0.4, 0.8)),
                 ggx((0.9, 0.8, 0.6), 0.2), ggx((0.7, 0.7, 0.75), 0.5),
                 dielectric((1, 1, 1), (1, 1, 1), roughness=0.05, ior=1.5),
                 dielectric((1, 1, 1), (0.9, 1, 0.9), roughness=0.15, ior=1.5)]
    u = g.random(len(idx))
    pick = g.integers(0, 4, len(idx))
    tri_mat = np.where(u < 0.5, pick, np.where(u < 0.8, 4 + (pick & 1), 6 + (pick & 1))).astype(np.uint32)
    return {"pos": pos, "idx": idx, "quads": [scenes.translate_quad((0.5, 1.3, 0.5), 0.5, facing_down=True)],
            "tri_material": tri_mat, "materials": materials, "quad_radiance": np.full((1, 3), 8.0, np.float32),
            "env": (0.25, 0.3, 0.35), "ntris": len(idx)}


def random_camera(g: np.random.Generator):
    """Eye at distance U[1.6, 2.4] from the soup's centre, elevation within +-60 degrees; looks at
    the centre (jittered by up to 0.1); vertical fov U[35, 55] degrees."""
    d = g.uniform(1.6, 2.4)
    az = g.uniform(0, 2 * math.pi)
    el = g.uniform(-math.pi / 3, math.pi / 3)
    c = np.array([0.5, 0.5, 0.5])
    eye = c + d * np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)])
    target = c + g.uniform(-0.1, 0.1, 3)
    return camera(tuple(eye), tuple(target), (0.0, 1.0, 0.0), float(g.uniform(35.0, 55.0)))


def request_trace(n_requests, n_scenes=16, rate=20.0, seed=scenes.SEED_SERVE + 1, depth=8,
                  resolutions=RESOLUTIONS, spps=SPPS):
    """Poisson arrivals at `rate` per second: list of dict(t, scene, cam, width, height, spp, depth, seed)."""
    g = scenes.rng(seed)
    t = np.cumsum(g.exponential(1.0 / rate, n_requests))
    out = []
    for i in range(n_requests):
        res = int(g.choice(resolutions))
        out.append({"t": float(t[i]), "scene": int(g.integers(0, n_scenes)), "cam": random_camera(g), "width": res,
                    "height": res, "spp": int(g.choice(spps)), "depth": depth, "seed": int(seed * 1000 + i)})
    return out


def build_scenes(ctx, n_scenes=16, scale=1.0, seed=scenes.SEED_SERVE, width=8):
    """Builds the served scenes' accelerators (host-side trees on `ctx`). Returns list of
    (accel, scene dict)."""
    from .accel import Obvh, Qbvh
    cls = Obvh if width == 8 else Qbvh
    sizes = np.maximum((scene_sizes(n_scenes, seed) * scale).astype(np.int64), 64)
    out = []
    for i in range(n_scenes):
        sc = soup_scene(int(sizes[i]), seed * 1000 + i)
        acc = cls(ctx).build(sc["pos"], sc["idx"], quads=sc["quads"])
        out.append((acc, sc))
    return out


def add_scenes(server: Server, built):
    return [server.add_scene(acc, sc["materials"], tri_material=sc["tri_material"], quad_radiance=sc["quad_radiance"],
                             env_radiance=sc["env"]) for acc, sc in built]


def run_served(server: Server, trace, scene_ids, time_scale=1.0, on_submit=None):
    """Submits the trace at its arrival times (scaled) and waits for every request. Returns
    (per-request dicts with latency_s, wall seconds)."""
    t0 = time.perf_counter()
    ids = []
    for r in trace:
        dt = r["t"] * time_scale - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        rid = server.submit(scene_ids[r["scene"]], r["cam"], r["width"], r["height"], r["spp"], r["depth"], r["seed"])
        ids.append(rid)
        if on_submit:
            on_submit(rid)
    out = []
    for rid in ids:
        out.append(server.wait(rid))
    wall = time.perf_counter() - t0
    for o in out:
        o["latency_s"] = o["t_done"] - o["t_submit"]
    return ids, out, wall


def run_one_at_a_time(pts, trace, time_scale=1.0, per_job=None, width=8):
    """The plain path: each request rendered alone (``PathTracing.render``, result copied to the
    host) in arrival order, starting when it has arrived and the previous one is done. With
    `per_job` = (ctx, built scenes), each request first builds and uploads its scene and frees it
    afterwards, as one SORT render job does (the scene is loaded and its accelerator built per
    job); otherwise `pts` holds every scene resident. `width`: node width of the per-job trees.
    Returns (latencies in seconds, wall seconds, ray totals)."""
    from .accel import Obvh, Qbvh
    from .integrator import PathTracing
    t0 = time.perf_counter()
    lat, ext, shadow = [], 0, 0
    for r in trace:
        arrive = r["t"] * time_scale
        now = time.perf_counter() - t0
        if now < arrive:
            time.sleep(arrive - now)
        if per_job:
            ctx, built = per_job
            sc = built[r["scene"]][1]
            acc = (Obvh if width == 8 else Qbvh)(ctx).build(sc["pos"], sc["idx"], quads=sc["quads"])
            pt = PathTracing(acc, sc["materials"], tri_material=sc["tri_material"], quad_radiance=sc["quad_radiance"],
                             env_radiance=sc["env"])
        else:
            pt = pts[r["scene"]]
        pt.render(r["cam"], r["width"], r["height"], r["spp"], seed=r["seed"], max_depth=r["depth"]).cpu()
        lat.append(time.perf_counter() - t0 - arrive)
        st = pt.stats()
        ext += st["extension_rays"]
        shadow += st["shadow_rays"]
        if per_job:
            pt.close()
            acc.release()
    return np.array(lat), time.perf_counter() - t0, (ext, shadow)


def choose_width(ctx, scale=1.0, n_sample=4, log=print):
    """Node width for the served scenes, picked by the autotuner (autotune.tune over the two
    widths, default knobs otherwise) on a sample of the served workload: one 512x512, 16-spp render
    (4M paths, enough to fill the GPU) on each of the `n_sample` largest scenes, trees of each
    width. Returns (width, {width: ms})."""
    from . import autotune
    from .integrator import PathTracing
    sizes = scene_sizes()
    pick = [int(i) for i in np.argsort(-sizes)[:n_sample]]
    g = np.random.default_rng(12345)
    cams = [random_camera(g) for _ in pick]
    runners, keep = {}, []
    for w in (8, 4):
        built = build_scenes(ctx, scale=scale, width=w)
        pts = [PathTracing(built[i][0], built[i][1]["materials"], tri_material=built[i][1]["tri_material"],
                           quad_radiance=built[i][1]["quad_radiance"], env_radiance=built[i][1]["env"]) for i in pick]
        keep.append((built, pts))

        def run(pts=pts):
            for pt, cam in zip(pts, cams):
                pt.render(cam, 512, 512, 16, seed=7)
        runners[w] = run
    best, results = autotune.tune(ctx, runners, lds=(0,), refills=(0,), blocks=(0,), leaf_batches=(0,), leaf_mins=(0,),
                                  reps=3)
    for built, pts in keep:
        for pt in pts:
            pt.close()
        for acc, _ in built:
            acc.release()
    ms = {int(r["width"]): round(r["ms"], 2) for r in results}
    log(f"[served] node width {best['width']} (sample render ms by width: {ms})")
    return int(best["width"]), ms


def _pct(x, q):
    return float(np.percentile(np.asarray(x), q))


def bench_served(ctx, n_requests=60, rate=20.0, scale=1.0, wave_paths=0, inflight_waves=0, graphs=False, width=0,
                 log=print):
    """Served workload, both ways, over trees of node width `width` (0: chosen by choose_width;
    both ways use the same trees). Returns a dict of p50/p99 latency and throughput for the
    Poisson-arrival run and for a saturated run (every request submitted at once), served versus
    one request at a time. Images are not compared here (the tests do that)."""
    from .integrator import PathTracing
    width_ms = None
    if width == 0:
        width, width_ms = choose_width(ctx, scale=scale, log=log)
    t = time.perf_counter()
    built = build_scenes(ctx, scale=scale, width=width)
    log(f"[served] built {len(built)} scenes ({sum(sc['ntris'] for _, sc in built)} tris) in "
        f"{time.perf_counter() - t:.1f}s")
    trace = request_trace(n_requests, rate=rate)
    paths = sum(r["width"] * r["height"] * r["spp"] for r in trace)
    srv = Server(ctx, wave_paths=wave_paths, inflight_waves=inflight_waves, graphs=graphs)
    ids = add_scenes(srv, built)
    # warm-up: load every scene and run one tiny request on each
    for i in ids:
        srv.wait(srv.submit(i, random_camera(np.random.default_rng(i)), 32, 32, 1))
    st0 = srv.stats()
    res = {}
    for name, ts in (("poisson", 1.0), ("saturated", 0.0)):
        s0 = srv.stats()
        _, out, wall = run_served(srv, trace, ids, time_scale=ts)
        s1 = srv.stats()
        lat = [o["latency_s"] for o in out]
        rays = (s1["extension_rays"] - s0["extension_rays"]) + (s1["shadow_rays"] - s0["shadow_rays"])
        res[name] = {"served": {"p50_ms": round(_pct(lat, 50) * 1e3, 1), "p99_ms": round(_pct(lat, 99) * 1e3, 1),
                                "wall_s": round(wall, 3), "mpaths_s": round(paths / wall / 1e6, 2),
                                "mrays_s": round(rays / wall / 1e6, 2), "waves": s1["waves"] - s0["waves"],
                                "device_busy_s": round((s1["device_busy_us"] - s0["device_busy_us"]) / 1e6, 3)}}
        log(f"[served] {name}: served {res[name]['served']}")
    info = srv.stats()
    srv.close()
    pts = [PathTracing(acc, sc["materials"], tri_material=sc["tri_material"], quad_radiance=sc["quad_radiance"],
                       env_radiance=sc["env"]) for acc, sc in built]
    # warm-up: each resident scene renders its largest request of the trace once, so no timed
    # one-at-a-time render grows its path workspace
    largest = {}
    for r in trace:
        k = r["width"] * r["height"] * r["spp"]
        if k > largest.get(r["scene"], (0, None))[0]:
            largest[r["scene"]] = (k, r)
    for i, pt in enumerate(pts):
        if i in largest:
            r = largest[i][1]
            pt.render(r["cam"], r["width"], r["height"], r["spp"], seed=r["seed"], max_depth=r["depth"]).cpu()
        else:
            pt.render(random_camera(np.random.default_rng(i)), 32, 32, 1).cpu()
    for name, ts in (("poisson", 1.0), ("saturated", 0.0)):
        for key, job in (("one_at_a_time", None), ("one_at_a_time_per_job", (ctx, built))):
            lat, wall, (ext, shadow) = run_one_at_a_time(pts, trace, time_scale=ts, per_job=job, width=width)
            res[name][key] = {"p50_ms": round(_pct(lat, 50) * 1e3, 1), "p99_ms": round(_pct(lat, 99) * 1e3, 1),
                              "wall_s": round(wall, 3), "mpaths_s": round(paths / wall / 1e6, 2),
                              "mrays_s": round((ext + shadow) / wall / 1e6, 2)}
            log(f"[served] {name}: {key.replace('_', '-')} {res[name][key]}")
        s = res[name]["served"]
        res[name]["throughput_ratio"] = round(res[name]["one_at_a_time"]["wall_s"] / s["wall_s"], 3)
        res[name]["throughput_ratio_per_job"] = round(res[name]["one_at_a_time_per_job"]["wall_s"] / s["wall_s"], 3)
    for pt in pts:
        pt.close()
    res["server"] = {"wave_paths": wave_paths or (1 << 25), "inflight_waves": inflight_waves or 3, "graphs": graphs,
                     "width": width, "width_sample_ms": width_ms}
    res["requests"] = n_requests
    res["rate_per_s"] = rate
    res["paths"] = paths
    res["scenes"] = len(built)
    res["scene_tris"] = [int(sc["ntris"]) for _, sc in built]
    res["resident_bytes"] = info["peak_bytes"]
    res["scene_loads"] = info["scene_loads"] - st0["scene_loads"]
    res["evictions"] = info["evictions"]
    return res
