"""Benchmark: incoherent closest-hit Mrays/s on the 1M-triangle intersection workload.

Workload (one GPU): a seeded 2^20-triangle soup (centroids U[0,1]^3, sizes LogNormal(ln 0.004,
0.5)), its SAH OBVH (max depth 28, max leaf 16) resident in HBM, and three 4,194,304-ray sets:
coherent 2048^2 pinhole rays, incoherent cosine-hemisphere bounce rays from random camera hits
(tmin 1e-4) and shadow segments to a 0.25^2 light quad (tmin 0.01, tmax len-0.01).

A step is one closest-hit traversal of the 4M incoherent rays (rays and BVH already resident).
value = incoherent Mrays/s over the timed steps. Coherent closest-hit and shadow any-hit rates
are timed the same way and reported beside it.

roofline: HBM-bound. Algorithmic bytes per ray = 32 (ray) + 16 (hit) + 224 * inner nodes
visited + 36 * triangles tested, with the visit counts from the traversal's own counters (equal
to the oracle's on these rays); achieved = bytes per launch / mean launch time (HIP events on
the launch stream). roofline.memory_pipeline adds the ceiling the kernel actually meets: lanes
read their own records, so compute units return 16-byte chunks at a bounded rate; a probe kernel
measures that rate for the traversal's two access patterns and gives the launch time it allows.

cpu_baseline: the oracle (C++ AVX restatement of SORT's Obvh traversal, std::thread pool, 4096-ray
chunks) on the same incoherent rays, on this host's cores.

Further legs, reported beside the headline under their own keys (--legs selects them):
  pt       -- the wavefront path tracer: the Cornell box (256^2, 16 spp, depth 5, OBVH) and the
              2M-triangle spheres scene (1024^2, 64 spp, depth 8, node width picked by the
              autotuner), refill threshold picked by the autotuner per frame (the default-knob
              rate beside it), samples/s and Mrays/s (median of 3 frames after one at full size),
              with the oracle's CPU path tracer timed on a bounded crop of each;
  batched  -- 64 scenes in one concatenated pool, one persistent launch over all their incoherent
              rays, against one launch per scene;
  served   -- the render server under Poisson arrivals (20 requests/s over 16 resident scenes) and
              saturated, p50/p99 latency and throughput against one request at a time; both use
              trees of the node width the autotuner picks on a sample of the served renders.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--width", type=int, default=8, choices=[4, 8])
    p.add_argument("--cpu-runs", type=int, default=5)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--tris", type=int, default=1 << 20)
    p.add_argument("--legs", default="intersection,attenuation,multi,film,pt,batched,served")
    p.add_argument("--served-requests", type=int, default=60)
    p.add_argument("--served-width", type=int, default=0, choices=[0, 4, 8],
                   help="node width of the served scenes' trees (0: the autotuner's pick on a sample of the renders)")
    p.add_argument("--shade-order", type=int, default=0, choices=[0, 1, 2],
                   help="path tracer: shade each bounce's hits in spatial order (0 default, 1 never, 2 always)")
    p.add_argument("--pt-no-tune", action="store_true",
                   help="pt leg: no autotuning (Cornell on its OBVH, spheres on the 4-wide tree, default knobs), "
                        "so that a profile holds only the measured frames")
    p.add_argument("--copy-probe", action="store_true",
                   help="end with three 256 MiB device copies (HBM counter calibration)")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = "profiles/k_trace_obvh_1m_soup_incoherent.md"


def profiled_traffic(a):
    """Memory-side bytes per incoherent launch, measured by FETCH_SIZE and WRITE_SIZE profiler
    passes over this workload and calibrated on a device copy of known size, as recorded in the
    profile summary; None when the workload differs from the profiled one."""
    if a.tris != 1 << 20 or a.width != 8:
        return None
    try:
        with open(os.path.join(ROOT, TRAFFIC_PROFILE)) as f:
            m = re.search(r"^traffic_bytes_per_launch: (\d+)$", f.read(), re.M)
    except OSError:
        return None
    return int(m.group(1)) if m else None


# Algorithmic bytes of k_shade (pt.hip): per shaded queue entry it reads the path's 64-byte record
# (throughput, radiance, two pending contributions; pixel, sample and request ids in their w
# lanes), the path id (4 B), the hit (16 B) and ray (32 B), three vertex indices (12 B), three
# shading normals (36 B), two positions for the edge tangent (24 B) and the triangle's material id
# (4 B) = 192 B, and writes the 64-byte record; each continuation ray it emits is 40 B (ray, path
# id, scene) and each shadow ray 40 B (ray, slot, scene). Records shared by many paths (materials,
# lights, requests) are cache-resident and not counted.
SHADE_READ_B, SHADE_WRITE_B, SHADE_RAY_B = 192, 64, 40


def shade_roofline(pt, cam, W, H, spp, depth, torch):
    """k_shade's HBM roofline for one frame: the frame rendered once more with stage timing (events
    around every bounce's shading launch on the stream it runs on), algorithmic bytes from the
    frame's ray counts (entries shaded = extension rays; continuations = extension rays - paths)."""
    pt.set_timing(True)
    pt.render(cam, W, H, spp, seed=1, max_depth=depth)
    torch.cuda.synchronize()
    st, tm = pt.stats(), pt.stage_times()
    pt.set_timing(False)
    ext, shadow, paths = st["extension_rays"], st["shadow_rays"], st["paths"]
    b = ext * (SHADE_READ_B + SHADE_WRITE_B) + (ext - paths) * SHADE_RAY_B + shadow * SHADE_RAY_B
    gbs = b / (tm["shade_ms"] * 1e-3) / 1e9 if tm["shade_ms"] > 0 else 0.0
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "kernel": "k_shade", "bytes": int(b),
            "entries": int(ext), "launches": tm["bounces"], "kernel_ms": round(tm["shade_ms"], 2),
            "extend_ms": round(tm["extend_ms"], 2), "shadow_ms": round(tm["shadow_ms"], 2),
            "note": "launch times on the render's overlapping internal streams (shared GPU time included)"}


def pt_tune(ctx, sc, cam, W, H, spp, depth, widths):
    """Node width and refill threshold for a path-traced frame, picked by the autotuner
    (autotune.tune over `widths` and refill 16 / 32 / 48, other knobs at their defaults) on a
    sample render of the frame: 1/16 of its samples, or the whole frame when that is under 2^20
    paths (the Cornell frame). The images do not depend on either knob. Returns (best tuning,
    {"w<width>/r<refill>": ms})."""
    from sort_amd import Obvh, Qbvh, autotune, integrator as I
    sspp = max(1, spp // 16)
    if W * H * sspp < (1 << 20):
        sspp = spp
    runners, keep = {}, []
    for w in widths:
        acc = (Obvh if w == 8 else Qbvh)(ctx).build(sc["pos"], sc["idx"], sc["quads"])
        pt = I.PathTracing(acc, sc["materials"], tri_material=sc.get("tri_material"), normals=sc.get("normals"),
                           quad_radiance=sc["quad_radiance"], env_radiance=sc["env"])
        keep.append((acc, pt))
        runners[w] = lambda pt=pt: pt.render(cam, W, H, sspp, seed=7, max_depth=depth)
    best, results = autotune.tune(ctx, runners, lds=(0,), refills=(16, 32, 48), blocks=(0,), leaf_batches=(0,),
                                  leaf_mins=(0,), reps=3)
    for acc, pt in keep:
        pt.close()
        acc.release()
    return best, {f"w{r['width']}/r{r['refill']}": round(r["ms"], 3) for r in results}


def leg_pt(ctx, cpu=True, tune=True):
    """Path tracer throughput on the Cornell box (OBVH, as its configuration names) and on the
    2M-triangle spheres scene (node width picked by the autotuner), with the refill threshold the
    autotuner picks for each frame (the rate with the default knobs is reported beside it). Each
    frame is rendered once at its full size before the timed frames, so the path workspace is
    sized outside the timing."""
    import torch

    from sort_amd import Obvh, Qbvh, batch, integrator as I, scenes
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    out = {}
    cases = [("cornell", scenes.cornell_sphere(), 256, 256, 16, 5, 3, (8,)),
             ("spheres", scenes.spheres_scene(), 1024, 1024, 64, 8, 3, (8, 4))]
    for name, sc, W, H, spp, depth, reps, widths in cases:
        cam = I.camera(**sc["camera"])
        if tune:
            best, tune_ms = pt_tune(ctx, sc, cam, W, H, spp, depth, widths)  # leaves `best` applied
        else:
            best, tune_ms = {"width": widths[-1], "refill": 16}, {}
        width = int(best["width"])
        t = time.perf_counter()
        acc = (Obvh if width == 8 else Qbvh)(ctx).build(sc["pos"], sc["idx"], sc["quads"])
        pt = I.PathTracing(acc, sc["materials"], tri_material=sc.get("tri_material"), normals=sc.get("normals"),
                           quad_radiance=sc["quad_radiance"], env_radiance=sc["env"])
        build_s = time.perf_counter() - t
        pt.render(cam, W, H, spp, seed=99, max_depth=depth)
        torch.cuda.synchronize()
        secs, ext, shadow = [], 0, 0
        for r in range(reps):
            t = time.perf_counter()
            pt.render(cam, W, H, spp, seed=1 + r, max_depth=depth)
            torch.cuda.synchronize()
            secs.append(time.perf_counter() - t)
            st = pt.stats()
            ext += st["extension_rays"]
            shadow += st["shadow_rays"]
        sec = float(np.median(secs))
        paths = W * H * spp
        res = {"size": f"{W}x{H}", "spp": spp, "depth": depth, "tris": int(len(sc["idx"])), "width": width,
               "ms": round(sec * 1e3, 2), "msamples_s": round(paths / sec / 1e6, 2),
               "mrays_s": round((ext + shadow) / reps / sec / 1e6, 2),
               "rays_per_path": round((ext + shadow) / reps / paths, 3), "build_s": round(build_s, 2)}
        res["tuning"] = {"width": width, "refill": int(best["refill"])}
        res["tune_sample_ms"] = tune_ms
        res["shade_roofline"] = shade_roofline(pt, cam, W, H, spp, depth, torch)
        batch.set_tuning(ctx)  # the same frame with the default knobs on the same tree
        dsecs = []
        for r in range(reps):
            t = time.perf_counter()
            pt.render(cam, W, H, spp, seed=1 + r, max_depth=depth)
            torch.cuda.synchronize()
            dsecs.append(time.perf_counter() - t)
        res["default_knobs_msamples_s"] = round(paths / float(np.median(dsecs)) / 1e6, 2)
        if cpu:
            from oracle.oracle import OraclePT
            o = OraclePT(sc)
            # the whole Cornell frame; for the spheres scene a band of rows through the middle of
            # the frame (ground, spheres and sky in the proportions of the full image's centre)
            rows = H if name == "cornell" else 64
            row0 = (H - rows) // 2
            ospp = spp if name == "cornell" else 4
            st = {}
            o.render(sc["camera"], W, H, ospp, depth, 1, threads=threads, crop=(W * row0, W * (row0 + rows)),
                     stats=st)
            cpaths = W * rows * ospp
            res["cpu_baseline"] = {"msamples_s": round(cpaths / st["seconds"] / 1e6, 3), "cores": threads,
                                   "kind": "port",
                                   "sample": f"rows {row0}..{row0 + rows - 1} of {W}x{H} at {ospp} spp, depth {depth}"}
            res["gpu_over_cpu"] = round(res["msamples_s"] / res["cpu_baseline"]["msamples_s"], 1)
        log(f"[pt] {name}: {res}")
        out[name] = res
        pt.close()
    return out


def leg_batched(ctx):
    """64 scenes, one pooled launch over all their rays versus one launch per scene."""
    import torch

    from sort_amd import autotune, batch
    t = time.perf_counter()
    w = batch.build_workload(ctx, n_scenes=64, width=8)
    build_s = time.perf_counter() - t
    pool = batch.Pool(ctx, w["accels"])
    d_rays = torch.from_numpy(w["rays"]).cuda()
    d_sid = torch.from_numpy(w["scene_of_ray"].view(np.int32)).cuda()
    out = torch.empty((len(w["rays"]), 4), dtype=torch.float32, device="cuda")

    def run_pool():
        pool.trace(d_rays, d_sid, out=out)

    def run_each():
        for i, acc in enumerate(w["accels"]):
            lo, hi = w["offsets"][i], w["offsets"][i + 1]
            acc.get_intersect(d_rays[lo:hi])

    n = len(w["rays"])
    ms_pool = autotune._time(run_pool, reps=5, warmup=2)
    ms_each = autotune._time(run_each, reps=5, warmup=2)
    res = {"scenes": 64, "tris": int(w["tris"].sum()), "rays": n, "pool_mrays_s": round(n / ms_pool / 1e3, 1),
           "per_scene_mrays_s": round(n / ms_each / 1e3, 1), "speedup": round(ms_each / ms_pool, 2),
           "pool_bytes": pool.info()["device_bytes"], "build_s": round(build_s, 1)}
    # the autotuner over node width (the same scenes as 4-wide trees), LDS stack depth, compaction
    # threshold (refill), resident workgroups per CU and leaf batch, timed on the whole workload
    from sort_amd import Qbvh
    t = time.perf_counter()
    pool4 = batch.Pool(ctx, [Qbvh(ctx).build(pos, idx) for pos, idx in w["geometry"]])
    build4_s = time.perf_counter() - t

    def runner(p):
        return lambda: p.trace(d_rays, d_sid, out=out)

    pools = {8: pool, 4: pool4}
    t = time.perf_counter()
    best, results = autotune.tune(ctx, {8: runner(pool), 4: runner(pool4)}, lds=(16, 8), refills=(8, 16, 32),
                                  blocks=(0, 2), leaf_batches=(16, 32), reps=3)
    tune_s = time.perf_counter() - t
    ms_tuned = autotune._time(runner(pools[best["width"]]), reps=5, warmup=2)
    batch.set_tuning(ctx)  # defaults for the legs that follow
    res["autotune"] = {"best": {k: best[k] for k in ("width", "lds_entries", "refill", "blocks_per_cu", "leaf_batch")},
                       "candidates": len(results), "tune_s": round(tune_s, 1), "build_4wide_s": round(build4_s, 1),
                       "tuned_mrays_s": round(n / ms_tuned / 1e3, 1),
                       "tuned_over_default": round(ms_pool / ms_tuned, 3)}
    log(f"[batched] {res}")
    pool.close()
    pool4.close()
    return res


def leg_attenuation(acc, d_sh, sh, pos, idx, stream, a, cpu=True):
    """Transparent shadows (Scene::GetAttenuation): the 4M area-light shadow segments through the
    soup with a seeded table in which half of the triangles' materials are transparent (colour in
    [0.2, 1)), the rest opaque. GPU rate over the whole batch; the oracle on 1M of the segments."""
    import torch
    from sort_amd._lib import check, lib as _l
    n = d_sh.shape[0]
    g = np.random.default_rng(2024)
    tr = np.zeros((len(idx), 4), np.float32)
    tr[:, :3] = g.uniform(0.2, 1.0, (len(idx), 3)).astype(np.float32)
    tr[:, 3] = (g.random(len(idx)) < 0.5).astype(np.float32)
    d_tr = torch.from_numpy(tr).to(d_sh.device)
    out = torch.empty((n, 4), dtype=torch.float32, device=d_sh.device)
    lib = _l()

    def run():
        check(lib.sortgpu_trace_attenuation(acc.handle, d_sh.data_ptr(), n, d_tr.data_ptr(), len(idx),
                                            out.data_ptr(), stream.cuda_stream))

    _, ms, _ = timed(run, max(5, a.steps // 5), 2, stream)
    res = out.cpu().numpy()
    rec = {"workload": f"{n} area-light shadow segments, 50% transparent materials",
           "mrays_s": round(n / (ms * 1e-3) / 1e6, 3), "launch_ms_mean": round(ms, 4),
           "surfaces_per_ray": round(float(res[:, 3].mean()), 4),
           "blocked_frac": round(float(np.all(res[:, :3] == 0, axis=1).mean()), 4)}
    if cpu:
        from oracle.oracle import OracleBvh
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        ob = OracleBvh(pos, idx, width=a.width)
        sample = np.ascontiguousarray(sh[: 1 << 20])
        ob.attenuation(sample[: 1 << 14], tr, threads=threads)  # warm-up
        t0 = time.perf_counter()
        ref = ob.attenuation(sample, tr, threads=threads)
        secs = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(len(sample) / secs / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                               "kind": "port", "sample": f"first {len(sample)} segments, one run after a warm-up"}
        rec["gpu_over_cpu"] = round(rec["mrays_s"] / rec["cpu_baseline"]["value"], 2)
        rec["sample_bitwise_equal"] = bool(np.array_equal(ref.view(np.uint32), res[: len(sample)].view(np.uint32)))
    return rec


def leg_multi(acc, inc, pos, idx, stream, a, cpu=True):
    """BSSRDF probes (Fbvh::GetIntersect with BSSRDFIntersections): segments of length 0.04 through
    the 4M incoherent rays' surface points along their directions (starting 0.02 before the point,
    window (0, 0.04]: the root test rejects rays that start inside the scene box with a negative
    tmin), four materials at random over the triangles, each probe looking for the material of the
    surface its segment meets first (found with a closest-hit query before the timed region). GPU
    rate over the whole batch; the oracle on 1M of the probes."""
    import torch
    from sort_amd._lib import check, lib as _l
    probes = inc.copy()
    probes[:, 0:3] = probes[:, 0:3] - np.float32(0.02) * probes[:, 4:7]
    probes[:, 3] = np.float32(0.0)
    probes[:, 7] = np.float32(0.04)
    n = len(probes)
    g = np.random.default_rng(2025)
    tri_mat = g.integers(0, 4, len(idx)).astype(np.uint32)
    first = acc.get_intersect(probes)[:, 1].copy().view(np.uint32)
    ray_mat = np.where(first != 0xFFFFFFFF, tri_mat[np.minimum(first, len(idx) - 1)], 0).astype(np.uint32)
    dev = torch.device("cuda:0")
    d_p = torch.from_numpy(probes).to(dev)
    d_rm = torch.from_numpy(ray_mat.view(np.int32)).to(dev)
    d_tm = torch.from_numpy(tri_mat.view(np.int32)).to(dev)
    hits = torch.empty((n, 4, 4), dtype=torch.float32, device=dev)
    cnt = torch.empty((n,), dtype=torch.int32, device=dev)
    lib = _l()

    def run():
        check(lib.sortgpu_trace_multi(acc.handle, d_p.data_ptr(), n, d_rm.data_ptr(), d_tm.data_ptr(), len(idx),
                                      hits.data_ptr(), cnt.data_ptr(), stream.cuda_stream))

    _, ms, _ = timed(run, max(5, a.steps // 5), 2, stream)
    c = cnt.cpu().numpy()
    rec = {"workload": f"{n} probe segments of length 0.04 through surface points, 4 materials",
           "mrays_s": round(n / (ms * 1e-3) / 1e6, 3), "launch_ms_mean": round(ms, 4),
           "records_per_ray": round(float(c.mean()), 4), "full_frac": round(float((c == 4).mean()), 4)}
    if cpu:
        from oracle.oracle import OracleBvh
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        ob = OracleBvh(pos, idx, width=a.width)
        m = 1 << 20
        ob.multi(probes[: 1 << 14], ray_mat[: 1 << 14], tri_mat, threads=threads)  # warm-up
        t0 = time.perf_counter()
        ref, ref_c = ob.multi(probes[:m], ray_mat[:m], tri_mat, threads=threads)
        secs = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(m / secs / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
                               "sample": f"first {m} probes, one run after a warm-up"}
        rec["gpu_over_cpu"] = round(rec["mrays_s"] / rec["cpu_baseline"]["value"], 2)
        h = hits[:m].cpu().numpy()
        rec["sample_bitwise_equal"] = bool(np.array_equal(ref.view(np.uint32), h.view(np.uint32)) and
                                           np.array_equal(ref_c, c[:m].view(np.uint32)))
    return rec


def leg_film(stream):
    """SORT's EXR output of a 3840 x 2160 film: from HBM (half conversion on the GPU, 6 bytes per
    pixel copied to the host) and from host memory (conversion on one host core)."""
    import torch
    from sort_amd import film
    w, h = 3840, 2160
    g = torch.Generator(device="cuda:0").manual_seed(3)
    d = torch.rand((h, w, 3), generator=g, device="cuda:0", dtype=torch.float32)
    host = d.cpu().numpy()
    film.encode_exr(d, stream=stream.cuda_stream)  # warm-up
    t0 = time.perf_counter()
    for _ in range(5):
        dev_bytes = film.encode_exr(d, stream=stream.cuda_stream)
    t_dev = (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    host_bytes = film.encode_exr(host)
    t_host = time.perf_counter() - t0
    return {"workload": f"EXR of a {w}x{h} film ({len(dev_bytes)} bytes)", "device_film_ms": round(t_dev * 1e3, 2),
            "host_film_ms": round(t_host * 1e3, 2), "bytes_equal": dev_bytes == host_bytes}


def host_cpu():
    """The host CPU the baselines ran on: lscpu's model name, sockets, cores per socket and threads
    per core, with the CPUs this process may run on."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except OSError:
        usable = os.cpu_count()
    return {"model": info.get("Model name", "unknown"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "cpus_online": info.get("CPU(s)"), "cpus_usable": usable}


def probe_lib():
    """The benchmark's vector-memory probe (measure/chunk_probe.hip, built by the package build)."""
    import ctypes

    from sort_amd import lib
    path = os.path.join(ROOT, "measure", "libchunk_probe.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    lib()  # the library loads torch's HIP runtime first; the probe shares it
    p = ctypes.CDLL(path)
    p.chunk_probe.restype = ctypes.c_int
    p.chunk_probe.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_void_p, ctypes.c_void_p]
    return p


def memory_pipeline(n, st, mean_ms, value, stream, width):
    """The ceiling the traversal runs against: each lane reads its own records (an 8-wide node as
    8 consecutive 16-byte chunks, 14 when it has more than four children; a 4-wide node as 7; a
    leaf record of two triangles, or one quad, as 5), so a launch is bound by the rate at which
    compute units return such chunks to lanes. Leaf records per ray are estimated as half the
    triangles tested plus a quarter of the leaves visited (a leaf with an odd count ends on a
    one-triangle record) plus the quads. The probe library replays those access patterns on an L2-resident buffer
    (the traversal's upper levels live in L2; its own L2 hit rate is lower, so this is an upper
    bound) and gives chunks/s; the ceiling is the launch time those rates allow for this launch's
    chunk counts."""
    import torch

    p = probe_lib()
    node_chunks = 14 if width == 8 else 7
    narrow_chunks = 8 if width == 8 else 7
    buf = torch.zeros(2 << 20, dtype=torch.uint8, device="cuda")  # 8,192 records of 256 B
    blocks = torch.cuda.get_device_properties(0).multi_processor_count * 12
    out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")
    iters = 256
    rates = {}
    for k in sorted({node_chunks, narrow_chunks, 5, 3}):
        def run():
            rc = p.chunk_probe(buf.data_ptr(), 8192, k, iters, blocks, out.data_ptr(), stream.cuda_stream)
            if rc != 0:
                raise RuntimeError(f"probe launch failed: {rc}")
        _, ms, _ = timed(run, 5, 2, stream)
        rates[k] = blocks * 256 * iters * k / (ms * 1e-3) / 1e9  # G chunks/s
    wide = st.get("wide_nodes", 0) if width == 8 else st["inner_nodes"]
    narrow = st["inner_nodes"] - wide
    records = st["tris"] / 2 + st["leaves"] / 4 + st["quads"]
    per_ray = {"wide_node": node_chunks * wide / n, "node": narrow_chunks * narrow / n,
               "leaf_record": 5 * records / n, "ray": 2.0}
    ns_per_ray = (per_ray["wide_node"] / rates[node_chunks] + per_ray["node"] / rates[narrow_chunks] +
                  per_ray["leaf_record"] / rates[5] + per_ray["ray"] / rates[3])
    ceiling = 1e3 / ns_per_ray  # Mrays/s
    chunks = sum(per_ray.values())
    return {"bound": "vector-memory return rate (16-byte chunks to lanes)",
            "chunks_per_ray": round(chunks, 1),
            "achieved_gchunks_s": round(chunks * n / (mean_ms * 1e-3) / 1e9, 1),
            "wide_node_frac": round(wide / max(st["inner_nodes"], 1), 4),
            "probe_gchunks_s": {f"{k}_per_lane": round(v, 1) for k, v in sorted(rates.items(), reverse=True)},
            "ceiling_mrays_s": round(ceiling, 1), "frac": round(value / ceiling, 4)}


def timed(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record(stream)
        fn()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    return wall, float(np.mean(per)), float(np.median(per))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                                   "(rays shard across GPUs with a replicated BVH)"}))
        sys.exit(2)
    import torch

    from sort_amd import Context, Obvh, Qbvh, scenes

    t_setup = time.perf_counter()
    pos, idx = scenes.soup(a.tris)
    ctx = Context(0)
    acc = (Obvh if a.width == 8 else Qbvh)(ctx).build(pos, idx)
    info = acc.info()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)

    cam = scenes.pinhole_rays(2048, 2048)
    d_cam = torch.from_numpy(cam).to(dev)
    cam_hits = acc.get_intersect(d_cam).cpu().numpy()
    inc = scenes.incoherent_rays(cam, cam_hits, pos, idx)
    sh = scenes.shadow_rays(cam, cam_hits, pos, idx)
    d_inc = torch.from_numpy(inc).to(dev)
    d_sh = torch.from_numpy(sh).to(dev)
    n = d_inc.shape[0]
    hits = torch.empty((n, 4), dtype=torch.float32, device=dev)
    occ = torch.empty((n,), dtype=torch.uint8, device=dev)
    from sort_amd._lib import check, lib as _l
    lib = _l()
    s_ptr = stream.cuda_stream

    def trace_inc():
        check(lib.sortgpu_trace_closest(acc.handle, d_inc.data_ptr(), n, hits.data_ptr(), s_ptr))

    def trace_cam():
        check(lib.sortgpu_trace_closest(acc.handle, d_cam.data_ptr(), n, hits.data_ptr(), s_ptr))

    def trace_sh():
        check(lib.sortgpu_trace_any(acc.handle, d_sh.data_ptr(), n, occ.data_ptr(), s_ptr))

    # visit counts for the bytes model (same traversal, counters on)
    _, st_inc = acc.counters(d_inc, 0)
    _, st_cam = acc.counters(d_cam, 0)
    _, st_sh = acc.counters(d_sh, 1)
    setup_s = time.perf_counter() - t_setup

    node_b = 224 if a.width == 8 else 112

    def bytes_per_ray(st, out_b):
        return 32 + out_b + node_b * st["inner_nodes"] / n + 36 * (st["tris"] + st["quads"]) / n

    # timed region: K steps of the incoherent closest-hit traversal
    torch.cuda.synchronize()
    wall, mean_ms, med_ms = timed(trace_inc, a.steps, a.warmup, stream)
    inc_mrays = n * a.steps / wall / 1e6
    _, cam_ms, _ = timed(trace_cam, max(5, a.steps // 5), 2, stream)
    _, sh_ms, _ = timed(trace_sh, max(5, a.steps // 5), 2, stream)
    att = None
    if "attenuation" in set(a.legs.split(",")):
        att = leg_attenuation(acc, d_sh, sh, pos, idx, stream, a, cpu=not a.no_cpu)
    multi = None
    if "multi" in set(a.legs.split(",")):
        multi = leg_multi(acc, inc, pos, idx, stream, a, cpu=not a.no_cpu)

    b_inc = bytes_per_ray(st_inc, 16)
    achieved = b_inc * n / (mean_ms * 1e-3) / 1e9

    cpu = None
    if not a.no_cpu:
        from oracle.oracle import OracleBvh
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        ob = OracleBvh(pos, idx, width=a.width)
        ob.closest(inc[: 1 << 16], threads=threads)  # warm-up
        secs = []
        for _ in range(a.cpu_runs):
            st = {}
            ob.closest(inc, threads=threads, stats=st)
            secs.append(st["seconds"])
        cpu_mrays = n / float(np.median(secs)) / 1e6
        cpu = {"value": round(cpu_mrays, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"all {n} incoherent rays, median of {a.cpu_runs} runs after a warm-up; oracle = "
                         f"C++ AVX (-mavx -mno-fma) restatement of SORT's Obvh traversal, {threads} threads",
               "host": host_cpu()}

    out = {
        "metric": "incoherent Mrays/s",
        "value": round(inc_mrays, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"1M-triangle seeded soup ({a.tris} tris), "
                               f"{'OBVH' if a.width == 8 else 'QBVH'} depth 28 leaf 16, "
                               "4,194,304 incoherent closest-hit rays per step",
                   "rays_per_step": n, "width": a.width,
                   "coherent_mrays_s": round(n / (cam_ms * 1e-3) / 1e6, 3),
                   "shadow_mrays_s": round(n / (sh_ms * 1e-3) / 1e6, 3),
                   "bvh_build_ms": info["build_ms"], "bvh_depth": info["depth"],
                   "setup_s": round(setup_s, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(a),
                     "traffic_source": TRAFFIC_PROFILE if profiled_traffic(a) else None,
                     "bytes_per_ray": round(b_inc, 1), "launch_ms_mean": round(mean_ms, 4),
                     "launch_ms_median": round(med_ms, 4),
                     "visits_per_ray": {"inner_nodes": round(st_inc["inner_nodes"] / n, 3),
                                        "tris": round(st_inc["tris"] / n, 3)},
                     "coherent": {"bytes_per_ray": round(bytes_per_ray(st_cam, 16), 1),
                                  "frac": round(bytes_per_ray(st_cam, 16) * n / (cam_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
                     "shadow": {"bytes_per_ray": round(bytes_per_ray(st_sh, 4), 1),
                                "frac": round(bytes_per_ray(st_sh, 4) * n / (sh_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["config"]["gpu_over_cpu"] = round(inc_mrays / cpu["value"], 2)
    out["roofline"]["memory_pipeline"] = memory_pipeline(n, st_inc, mean_ms, inc_mrays, stream, a.width)
    if att:
        out["attenuation"] = att
    if multi:
        out["multi"] = multi
    legs = set(a.legs.split(","))
    if "film" in legs:
        out["film"] = leg_film(stream)
    del d_inc, d_sh, d_cam, hits, occ
    acc.release()
    torch.cuda.empty_cache()
    from sort_amd.batch import set_tuning
    set_tuning(ctx, shade_order=a.shade_order)
    if "pt" in legs:
        out["pt"] = leg_pt(ctx, cpu=not a.no_cpu, tune=not a.pt_no_tune)
    if "batched" in legs:
        out["batched"] = leg_batched(ctx)
        set_tuning(ctx, shade_order=a.shade_order)  # the batched leg leaves the defaults applied
    if "served" in legs:
        from sort_amd.serve import bench_served, choose_width
        width, width_ms = (a.served_width, None) if a.served_width else choose_width(ctx, log=log)
        set_tuning(ctx, shade_order=a.shade_order)  # the width choice leaves the defaults applied
        out["served"] = bench_served(ctx, n_requests=a.served_requests, width=width, log=log)
        out["served"]["server"]["width_sample_ms"] = width_ms
    if a.copy_probe:
        # a device-to-device copy of known size: its HBM counters per dispatch calibrate the
        # traversal's (FETCH_SIZE / WRITE_SIZE runs of a profiler over this program)
        x = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = x.numel() * 4
    print(json.dumps(out))


if __name__ == "__main__":
    main()
