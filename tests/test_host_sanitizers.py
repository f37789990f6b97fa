"""The host side of the library (planner, plan cache, ragged planning, autotuner, VJP front-end)
built with AddressSanitizer and UndefinedBehaviorSanitizer on the host pass, and the C-ABI tests
run against that build in a child process with the ASan runtime preloaded.  Without a GPU the
calls stop at the argument checks and the host-only planning (afft_describe)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "accelerate-fft_amd"))


def sanitized_env():
    import build
    rt = build.san_runtime()
    if rt is None:
        pytest.skip("no clang ASan runtime in this image")
    lib = build.SAN_LIB
    if build.sanitized_stale():
        lib = build.build_sanitized(verbose=False)
    env = dict(os.environ)
    env.update({"AFFT_LIB": lib, "LD_PRELOAD": rt,
                # the interpreter itself is not instrumented and keeps its allocations to the end
                "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1",
                "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})
    return env


def test_abi_under_asan_and_ubsan():
    env = sanitized_env()
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_abi.py")], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "passed" in r.stdout


def test_sanitized_build_is_instrumented():
    env = sanitized_env()
    out = subprocess.run(["nm", "-D", env["AFFT_LIB"]], capture_output=True, text=True).stdout
    assert "__asan_report" in out and "__ubsan_handle" in out
