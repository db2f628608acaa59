"""Serving engine on the host-only stand-in device: scheduling, admission under the HBM budget,
LRU eviction, cancellation, timeouts and injected failures, with results compared alone versus in
the mix. The stand-in's films hold a hash of each path key, accumulated in sample order like the
device films, so a result is independent of its wave placement exactly when the scheduler keeps
whole samples in order."""
import threading

import numpy as np
import pytest

from sort_amd import Context, SortGpuError, scenes
from sort_amd.integrator import camera, lambert
from sort_amd.serve import Server


@pytest.fixture(scope="module")
def hctx():
    return Context(-1)


@pytest.fixture(scope="module")
def built(hctx):
    from sort_amd.accel import Obvh
    out = []
    for i, n in enumerate([300, 2000, 800, 5000]):
        pos, idx = scenes.soup(n, seed=77 + i)
        out.append(Obvh(hctx).build(pos, idx))
    return out


CAM = camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5), (0, 1, 0), 40.0)


def make_server(hctx, built, **kw):
    s = Server(hctx, host_device=True, **kw)
    ids = [s.add_scene(a, [lambert((0.7, 0.7, 0.7))]) for a in built]
    return s, ids


REQS = [(0, 16, 16, 4), (1, 32, 32, 16), (2, 24, 20, 7), (3, 64, 64, 3), (1, 8, 8, 64), (0, 40, 40, 9)]


def alone(hctx, built, sc, w, h, spp, seed, **kw):
    s, ids = make_server(hctx, built, manual=True, **kw)
    rid = s.submit(ids[sc], CAM, w, h, spp, seed=seed)
    s.drain()
    assert s.info(rid)["state"] == "done"
    r = s.result(rid)
    s.close()
    return r


def test_alone_equals_in_mix_manual(hctx, built):
    s, ids = make_server(hctx, built, manual=True, wave_paths=5000, max_active=4)
    rids = [s.submit(ids[sc], CAM, w, h, spp, seed=100 + k) for k, (sc, w, h, spp) in enumerate(REQS)]
    s.drain()
    st = s.stats()
    assert st["done"] == len(REQS) and st["failed"] == 0
    assert st["waves"] > len(REQS)  # samples of different requests shared waves
    for k, (sc, w, h, spp) in enumerate(REQS):
        got = s.result(rids[k])
        ref = alone(hctx, built, sc, w, h, spp, 100 + k)
        assert got.shape == (h, w, 3)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), k
    s.close()


def test_threaded_concurrent_submitters(hctx, built):
    s, ids = make_server(hctx, built, wave_paths=4096)
    results = {}

    def client(k):
        sc, w, h, spp = REQS[k % len(REQS)]
        rid = s.submit(ids[sc], CAM, w, h, spp, seed=500 + k)
        info = s.wait(rid, 30.0)
        assert info["state"] == "done", info
        results[k] = (s.result(rid), (sc, w, h, spp))

    th = [threading.Thread(target=client, args=(k,)) for k in range(12)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert len(results) == 12
    for k, (img, (sc, w, h, spp)) in results.items():
        ref = alone(hctx, built, sc, w, h, spp, 500 + k)
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), k
    s.close()


def test_budget_eviction_lru(hctx, built):
    probe, ids = make_server(hctx, built, manual=True, wave_paths=4096)
    fixed = probe.stats()["workspace_bytes"]
    sizes = []
    for i in ids:  # a resident scene's bytes = bytes in use after its request is done, less the workspace
        before = probe.stats()["bytes_in_use"]
        probe.submit(i, CAM, 8, 8, 1)
        probe.drain()
        sizes.append(probe.stats()["bytes_in_use"] - before)
    probe.close()
    assert min(sizes) > 0
    # room for the workspace, the two largest scenes and two films, never for all four scenes
    two = sum(sorted(sizes)[-2:])
    s, ids = make_server(hctx, built, manual=True, wave_paths=4096, max_active=2,
                         hbm_budget=fixed + two + 2 * 16 * 16 * 16)
    order = [0, 1, 2, 3, 0, 3, 1, 2]
    rids = [s.submit(ids[sc], CAM, 16, 16, 5, seed=900 + k) for k, sc in enumerate(order)]
    s.drain()
    st = s.stats()
    assert st["done"] == len(order), [s.message(r) for r in rids]
    assert st["evictions"] > 0 and st["scene_loads"] > 4
    assert st["peak_bytes"] <= st["budget"]
    for k, sc in enumerate(order):
        ref = alone(hctx, built, sc, 16, 16, 5, 900 + k)
        assert np.array_equal(s.result(rids[k]).view(np.uint32), ref.view(np.uint32))
    s.close()


def test_request_over_budget_fails_cleanly(hctx, built):
    probe, _ = make_server(hctx, built, manual=True, wave_paths=1 << 16)
    fixed = probe.stats()["workspace_bytes"]
    probe.close()
    s, ids = make_server(hctx, built, manual=True, wave_paths=1 << 16, hbm_budget=fixed + 600_000)
    big = s.submit(ids[3], CAM, 256, 256, 2, seed=1)  # film of 1 MiB cannot fit
    ok = s.submit(ids[0], CAM, 16, 16, 2, seed=2)
    s.drain()
    assert s.info(big)["state"] == "failed" and s.info(big)["status"] == -4
    assert "budget" in s.message(big)
    assert s.info(ok)["state"] == "done"
    s.close()


@pytest.mark.parametrize("fail_at", [1, 2, 3])
def test_injected_allocation_failure(hctx, built, fail_at):
    # allocations in order: scene 0 load, film A, film B (scene 0 already resident) ...
    s, ids = make_server(hctx, built, manual=True, wave_paths=4096, fail_alloc_at=fail_at)
    a = s.submit(ids[0], CAM, 16, 16, 4, seed=11)
    b = s.submit(ids[0], CAM, 16, 16, 4, seed=12)
    c = s.submit(ids[2], CAM, 8, 8, 4, seed=13)
    s.drain()
    states = [s.info(r)["state"] for r in (a, b, c)]
    failed = {1: [0], 2: [0], 3: [1]}[fail_at]
    for k, r in enumerate((a, b, c)):
        if k in failed:
            assert states[k] == "failed" and s.info(r)["status"] == -3, (fail_at, states)
            assert "injected" in s.message(r)
        else:
            assert states[k] == "done", (fail_at, states)
            ref = alone(hctx, built, [0, 0, 2][k], [16, 16, 8][k], [16, 16, 8][k], 4, [11, 12, 13][k])
            assert np.array_equal(s.result(r).view(np.uint32), ref.view(np.uint32))
    s.close()


def test_cancel_and_timeout(hctx, built):
    s, ids = make_server(hctx, built, manual=True, wave_paths=1024, max_active=2)
    a = s.submit(ids[1], CAM, 32, 32, 40, seed=21)   # many waves
    b = s.submit(ids[0], CAM, 16, 16, 8, seed=22)
    c = s.submit(ids[2], CAM, 16, 16, 8, seed=23, timeout_s=1e-9)
    d = s.submit(ids[3], CAM, 16, 16, 8, seed=24)    # queued behind max_active
    s.cancel(d)
    s.step()
    s.step()
    s.cancel(a)  # active, with samples in flight
    s.drain()
    assert s.info(a)["state"] == "cancelled"
    assert s.info(c)["state"] == "timed_out"
    assert s.info(d)["state"] == "cancelled"
    assert s.info(b)["state"] == "done"
    ref = alone(hctx, built, 0, 16, 16, 8, 22)
    assert np.array_equal(s.result(b).view(np.uint32), ref.view(np.uint32))
    with pytest.raises(SortGpuError):
        s.result(a)
    st = s.stats()
    assert st["cancelled"] == 2 and st["timed_out"] == 1 and st["done"] == 1
    # every film is back: only the workspace and resident scenes stay accounted
    s.release(a)
    with pytest.raises(SortGpuError):
        s.info(a)
    s.close()


def test_invalid_requests_refused(hctx, built):
    s, ids = make_server(hctx, built, manual=True, wave_paths=4096)
    bad = [dict(scene=99), dict(width=0), dict(spp=0), dict(width=128, height=128),  # 16384 > wave
           dict(cam=camera((0.5, 0.5, 0.5), (0.5, 0.5, 0.5))), dict(cam=camera((0, 0, 0), (0, 1, 0), (0, 1, 0)))]
    for b in bad:
        kw = dict(scene=ids[0], cam=CAM, width=8, height=8, spp=1)
        kw.update(b)
        with pytest.raises(SortGpuError) as e:
            s.submit(kw["scene"], kw["cam"], kw["width"], kw["height"], kw["spp"])
        assert e.value.code == -1, b
    assert s.stats()["done"] == 0
    s.close()


def test_step_refused_in_threaded_mode(hctx, built):
    s, _ = make_server(hctx, built)
    with pytest.raises(SortGpuError):
        s.step()
    s.close()


def test_destroy_with_outstanding_requests(hctx, built):
    s, ids = make_server(hctx, built, wave_paths=256)
    for k in range(8):
        s.submit(ids[k % 4], CAM, 16, 16, 50, seed=k)
    s.close()  # cancels and waits for the waves in flight
