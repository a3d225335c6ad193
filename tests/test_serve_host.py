"""The serving loop's host logic (par-cd_amd/host/serve.cpp) under the
sanitizers, on the CPU.  tests/cpp/serve_stub.cpp links serve.cpp against a
stub of the device API that enforces the C ABI's threading contract (one
thread per context and per batch at a time) and computes a deterministic
result per request; the driver submits a staggered mixed stream from two
threads, checks every result against the request alone, exercises the error
path and shuts down with work still queued.

ThreadSanitizer is clang's (ROCm's LLVM): the libtsan of g++ 11 does not
intercept pthread_cond_clockwait, which std::condition_variable::wait_for
uses, and reports the waits as double locks and races."""
import os
import subprocess

import pytest

from conftest import HERE, ROOT

SRC = [os.path.join(HERE, "cpp", "serve_stub.cpp"), os.path.join(ROOT, "par-cd_amd", "host", "serve.cpp"),
       os.path.join(ROOT, "par-cd_amd", "host", "hostutil.cpp")]
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "par-cd_amd", "host"),
       "-I" + os.path.join(ROOT, "par-cd_amd", "csrc")]
CLANG = "/opt/rocm/llvm/bin/clang++"


def _build_run(tmp_path, cxx, san, env_extra):
    if not os.path.exists(cxx) and cxx != "g++":
        pytest.skip(cxx + " not present")
    exe = str(tmp_path / ("serve_stub_" + san.split(",")[0]))
    r = subprocess.run([cxx, "-std=c++17", "-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=" + san] + INC + SRC +
                       ["-lpthread", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    env = dict(os.environ, **env_extra)
    r = subprocess.run([exe, "400"], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert "differing 0" in out and "contract violations 0" in out, out[-1000:]
    return out


def test_server_under_thread_sanitizer(tmp_path):
    out = _build_run(tmp_path, CLANG, "thread", {"TSAN_OPTIONS": "halt_on_error=1"})
    assert "ThreadSanitizer" not in out


def test_server_under_address_and_undefined_sanitizers(tmp_path):
    out = _build_run(tmp_path, "g++", "address,undefined",
                     {"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})
    assert "AddressSanitizer" not in out and "runtime error" not in out
