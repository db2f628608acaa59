"""Build recipe for libsortgpu.so (HIP for gfx950) — run by ``__graft_entry__.build()``.

Every translation unit is compiled with the flags the parity contract needs:
  -ffp-contract=off                     no FMA contraction (the reference never fuses)
  -fno-gpu-flush-denormals-to-zero      keep fp32 denormals like x86 SSE/AVX
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE 1/x and a/b
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed fp32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_obj")
LIB = os.path.join(PKG, "libsortgpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-pthread",
    "-ffp-contract=off",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]

HOST_SOURCES = ["api.cpp", "accel_device.cpp", "bvh_build.cpp", "pool.cpp", "serve_engine.cpp", "serve_api.cpp",
                "scene_file.cpp", "hit_surface.cpp"]
# compiled as HIP: kernels, and host code that shares device record types with them
DEVICE_SOURCES = ["trace.hip", "order.hip", "multihit.hip", "film_out.hip", "pt.hip", "pt_host.cpp", "pt_api.cpp", "serve_device.cpp"]
# the vector-memory probe bench.py measures the traversal's ceiling with (benchmark code, its own
# library beside its source; not part of libsortgpu)
PROBE_SRC = os.path.join(ROOT, "measure", "chunk_probe.hip")
PROBE_LIB = os.path.join(ROOT, "measure", "libchunk_probe.so")


def _compile(src: str) -> str:
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON
    if src in DEVICE_SOURCES or src == PROBE_SRC:
        cmd += DEVICE + ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def _link(objs: list[str], out: str) -> None:
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    """Compiles every source and links ``sort_amd/libsortgpu.so`` (and the benchmark's probe
    ``measure/libchunk_probe.so``); returns the library's path."""
    srcs = HOST_SOURCES + DEVICE_SOURCES + [PROBE_SRC]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = dict(zip(srcs, ex.map(_compile, srcs)))
    _link([objs[s] for s in HOST_SOURCES + DEVICE_SOURCES], LIB)
    _link([objs[PROBE_SRC]], PROBE_LIB)
    if verbose:
        print(f"built {LIB} and {PROBE_LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
