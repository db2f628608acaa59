"""The serving engine under ThreadSanitizer and AddressSanitizer + UndefinedBehaviorSanitizer.

The engine (sort_amd/csrc/serve_engine.cpp) is plain C++ behind its device interface, so it is
compiled here with g++ together with tests/native/serve_engine_stress.cpp: six client threads
submit, cancel, time out, wait for and read ~240 requests while waves complete on a device thread
and a scripted share of film allocations fail. The driver checks every finished image against the
per-path values summed in sample order; the sanitizers check the concurrency and memory use."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "native", "serve_engine_stress.cpp"),
       os.path.join(ROOT, "sort_amd", "csrc", "serve_engine.cpp")]
INC = ["-I", os.path.join(ROOT, "sort_amd", "csrc"), "-I", os.path.join(ROOT, "include")]


# ThreadSanitizer from ROCm's clang: GCC 11's runtime does not intercept pthread_cond_clockwait
# (used by libstdc++'s steady-clock condition waits) and reports false double locks.
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.parametrize("compiler,flags", [(CLANG, ["-fsanitize=thread"]),
                                            ("g++", ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"])],
                         ids=["tsan", "asan_ubsan"])
def test_engine_under_sanitizer(tmp_path, compiler, flags):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not available")
    exe = str(tmp_path / "stress")
    cmd = [compiler, "-std=c++17", "-O1", "-g", "-pthread"] + flags + INC + SRC + ["-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout)
    assert r.returncode == 0, (r.stdout + r.stderr)[-6000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
