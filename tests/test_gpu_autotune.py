"""The batched form's autotuner (sort_amd/autotune.py) on the scaled 64-scene workload: it times
every candidate (node width 8/4, LDS stack depth, compaction threshold, resident workgroups, leaf
batch), leaves the fastest applied to the context, every candidate gives the same hits as the
default knobs, and the cache file returns the same choice without timing again."""
import json

import numpy as np
import pytest
import torch

from sort_amd import Qbvh, autotune, batch

pytestmark = pytest.mark.gpu


def test_autotune_picks_applies_and_caches(gpu_ctx, tmp_path):
    w8 = batch.build_workload(gpu_ctx, n_scenes=16, width=8, scale=0.05)
    acc4 = [Qbvh(gpu_ctx).build(pos, idx) for pos, idx in w8["geometry"]]  # the same scenes, 4-wide
    d_rays = torch.from_numpy(w8["rays"]).cuda()
    d_sid = torch.from_numpy(w8["scene_of_ray"].view(np.int32)).cuda()
    pools = {8: batch.Pool(gpu_ctx, w8["accels"]), 4: batch.Pool(gpu_ctx, acc4)}
    ref = pools[8].trace(d_rays, d_sid).cpu().numpy()
    seen = {}

    def runner(width):
        def run():
            out = pools[width].trace(d_rays, d_sid)
            seen.setdefault(width, []).append(out)
        return run

    cache = str(tmp_path / "tune.json")
    print(f"[autotune] {len(w8['rays'])} rays over 16 scenes, two widths", flush=True)
    best, results = autotune.tune(gpu_ctx, {8: runner(8), 4: runner(4)}, lds=(16, 8), refills=(8, 16, 32),
                                  blocks=(0, 2), leaf_batches=(16, 32), reps=2, cache=cache, signature="w16s05")
    try:
        assert len(results) == 2 * 2 * 3 * 2 * 2
        assert best["ms"] == min(r["ms"] for r in results) and best["width"] in (4, 8)
        # every candidate's hits are the default's, bit for bit
        for width, outs in seen.items():
            for o in outs:
                assert np.array_equal(o.cpu().numpy().view(np.uint32), ref.view(np.uint32)), width
        # the best stays applied: a trace now runs with it and still gives the same hits
        assert np.array_equal(pools[best["width"]].trace(d_rays, d_sid).cpu().numpy().view(np.uint32),
                              ref.view(np.uint32))
        with open(cache) as f:
            assert json.load(f)["w16s05"] == best
        again, res2 = autotune.tune(gpu_ctx, {8: runner(8)}, cache=cache, signature="w16s05")
        assert again == best and res2 == []
        print(f"[autotune] best {best}", flush=True)
    finally:
        batch.set_tuning(gpu_ctx)  # defaults for the tests that follow
        for p in pools.values():
            p.close()
