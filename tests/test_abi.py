"""C-ABI checks that need no GPU: the library loads, exports every declared symbol, rejects
bad arguments with status codes, and its host-side build equals the oracle's tree."""
import ctypes
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleBvh
from sort_amd import Context, Obvh, Qbvh, SortGpuError, header_symbols, lib, scenes
from sort_amd._lib import LIB_PATH


def test_library_exports_every_declared_symbol():
    l = lib()
    syms = header_symbols()
    assert len(syms) >= 13
    for s in syms:
        assert hasattr(l, s), s
    # the library exports nothing beyond the header (no measurement or test hooks)
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = sorted({ln.split()[-1] for ln in out.splitlines() if ln.split()[-1].startswith("sortgpu_")})
    assert exported == syms


def test_benchmark_probe_is_separate():
    """The benchmark's vector-memory probe is its own library beside bench.py, not part of the
    drop-in library or its header."""
    import bench
    p = bench.probe_lib()
    assert p.chunk_probe(None, 1, 3, 1, 1, None, None) != 0  # refused before any launch
    assert not hasattr(lib(), "chunk_probe")


def test_host_only_context_and_errors(host_ctx):
    from sort_amd._lib import check
    pos, idx = scenes.soup(64, seed=1)
    acc = Obvh(host_ctx).build(pos, idx)
    assert acc.valid
    # queries on a host-only context are refused with SORTGPU_ERR_STATE, without touching memory
    with pytest.raises(SortGpuError) as e:
        check(lib().sortgpu_trace_closest(acc.handle, ctypes.c_void_p(16), 1, ctypes.c_void_p(16), None),
              host_ctx.handle)
    assert e.value.code == -5
    bad = idx.copy()
    bad[0, 0] = len(pos)  # out-of-range vertex index
    with pytest.raises(SortGpuError) as e:
        Obvh(host_ctx).build(pos, bad)
    assert e.value.code == -1
    odd = Obvh(host_ctx)
    odd.width = 6  # neither OBVH nor QBVH
    with pytest.raises(SortGpuError) as e:
        odd.build(pos, idx)
    assert e.value.code == -1


def test_null_arguments_return_status():
    l = lib()
    assert l.sortgpu_ctx_create(0, None) == -1
    assert l.sortgpu_accel_build(None, None, 0, None, 0, None, 0, None, None) == -1
    assert l.sortgpu_accel_bbox(None, None) == -1
    assert l.sortgpu_accel_info(None, None) == -1
    assert l.sortgpu_trace_closest(None, None, 0, None, None) == -1


@pytest.mark.parametrize("width,cls", [(8, Obvh), (4, Qbvh)])
@pytest.mark.parametrize("n", [1, 17, 5000])
def test_host_build_equals_oracle_tree(host_ctx, width, cls, n):
    pos, idx = scenes.soup(n, seed=100 + n)
    acc = cls(host_ctx).build(pos, idx)
    ob = OracleBvh(pos, idx, width=width)
    assert np.array_equal(acc.prim_order(), ob.prim_order())
    assert np.array_equal(acc.tree_dump(), ob.tree_dump())
    assert np.array_equal(acc.bbox().view(np.uint32), ob.bbox().view(np.uint32))
    info = acc.info()
    assert info["depth"] == ob.info()["depth"] and info["leaves"] == ob.info()["leaves"]
    assert info["max_stack"] == (width - 1) * (info["depth"] - 1) + 1


def test_host_build_is_thread_count_independent(host_ctx):
    pos, idx = scenes.soup(40000, seed=3)
    a = Obvh(host_ctx, build_threads=1).build(pos, idx)
    b = Obvh(host_ctx, build_threads=8).build(pos, idx)
    assert np.array_equal(a.prim_order(), b.prim_order())
    assert np.array_equal(a.tree_dump(), b.tree_dump())


def test_host_build_with_quads_and_options(host_ctx):
    sc = scenes.cornell_sphere()
    for kw in [dict(), dict(max_node_depth=2), dict(max_pri_in_leaf=1), dict(max_node_depth=16, max_pri_in_leaf=8)]:
        acc = Obvh(host_ctx, **kw).build(sc["pos"], sc["idx"], sc["quads"])
        ob = OracleBvh(sc["pos"], sc["idx"], sc["quads"], max_depth=kw.get("max_node_depth", 28),
                       max_leaf=kw.get("max_pri_in_leaf", 16))
        assert np.array_equal(acc.prim_order(), ob.prim_order()), kw
        assert np.array_equal(acc.tree_dump(), ob.tree_dump()), kw


def test_context_outlives_objects_made_from_it():
    """Destroying a context first is safe: accelerators and servers made from it hold a reference,
    and the context is freed with the last of them."""
    from sort_amd import Context, Obvh, scenes
    from sort_amd.integrator import camera, lambert
    from sort_amd.serve import Server
    ctx = Context(-1)
    pos, idx = scenes.soup(500, seed=3)
    acc = Obvh(ctx).build(pos, idx)
    srv = Server(ctx, host_device=True, manual=True, wave_paths=4096)
    sid = srv.add_scene(acc, [lambert((0.5, 0.5, 0.5))])
    ctx.close()  # the caller's reference goes first
    cam = camera((0.5, 0.5, -1.5), (0.5, 0.5, 0.5))
    rid = srv.submit(sid, cam, 8, 8, 2)
    srv.drain()
    assert srv.info(rid)["state"] == "done"
    assert acc.info()["prims"] == 500
    srv.close()
    acc.release()
