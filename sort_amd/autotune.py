"""Autotuner for the persistent traversal kernels.

Searches node width (OBVH 8 / QBVH 4), LDS stack depth (8 / 16 entries per lane), refill chunk
(idle lanes that trigger taking new rays), resident workgroups per CU (the occupancy limit or 2)
and the two phase knobs of the traversal loop (lanes holding a leaf that end a node phase; lanes
below which a leaf phase ends), timing each candidate on a sample of the workload with HIP events,
and applies the fastest to the context. Every candidate gives the same hits; only speed differs.
Results can be cached in a JSON file keyed by a workload signature.
"""
from __future__ import annotations

import itertools
import json
import os

import numpy as np

from .batch import set_tuning

KEYS = ("refill", "lds_entries", "blocks_per_cu", "leaf_batch", "leaf_min")


def _time(fn, reps=3, warmup=1):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        ms.append(a.elapsed_time(b))
    return float(np.median(ms))


def _apply(ctx, t):
    set_tuning(ctx, *[int(t.get(k, 0)) for k in KEYS])


def tune(ctx, runners, lds=(16, 8), refills=(8, 16, 32), blocks=(0, 2), leaf_batches=(0,), leaf_mins=(0,), reps=3,
         cache=None, signature=None):
    """runners: {width: callable()} each tracing the sample with a tree of that width.
    Returns (best dict, all results); the best tuning stays applied to `ctx`."""
    if cache and signature and os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
        if signature in db:
            best = db[signature]
            _apply(ctx, best)
            return best, []
    results = []
    for width, run in runners.items():
        for l, r, b, lb, lm in itertools.product(lds, refills, blocks, leaf_batches, leaf_mins):
            t = {"refill": r, "lds_entries": l, "blocks_per_cu": b, "leaf_batch": lb, "leaf_min": lm}
            _apply(ctx, t)
            ms = _time(run, reps=reps)
            results.append(dict(t, width=width, ms=ms))
    best = min(results, key=lambda x: x["ms"])
    _apply(ctx, best)
    if cache and signature:
        db = {}
        if os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
        db[signature] = best
        with open(cache, "w") as f:
            json.dump(db, f, indent=1)
    return best, results
