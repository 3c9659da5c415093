"""CPU harness of the double-buffered stream path (pairhmm_stream_* and its Python driver).

The engine's C-ABI host code (pairhmm_capi.cpp, pairhmm_stream.cpp) is built with g++,
AddressSanitizer and UndefinedBehaviorSanitizer against a host-only HIP stand-in
(tests/stream_harness): device memory is malloc'd host memory, copies are memcpy, and a host
stand-in of the forward kernel writes a recomputable value for every pair.  The real Python wrappers
drive it in a subprocess with the sanitizer runtime preloaded: ``RegionStream`` with empty chunks
between real ones and with only the stream holding the output arrays (a use-after-free the
sanitizer reports at pairhmm_stream.cpp's finalize if an output array is released while its chunk is
in flight), then ``run_streaming`` with its spawn-context generator pool.
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "stream_harness")


def _asan_runtime():
    if shutil.which("g++") is None:
        return None
    p = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def harness_lib(tmp_path_factory):
    if _asan_runtime() is None:
        pytest.skip("g++ with AddressSanitizer is not available")
    out = str(tmp_path_factory.mktemp("stream_harness") / "libstream_harness.so")
    subprocess.run(["make", "-s", "-C", HARNESS, f"OUT={out}"], check=True)
    return out


def test_stream_path_under_address_sanitizer(harness_lib, tmp_path):
    env = dict(os.environ)
    env["LD_PRELOAD"] = _asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    p = subprocess.run([sys.executable, os.path.join(HARNESS, "drive_stream.py"), harness_lib, "600", "200", "3"],
                       capture_output=True, text=True, timeout=600, env=env, cwd=str(tmp_path))
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-6000:]
    assert "harness: done" in p.stdout
    assert "ERROR: AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr
