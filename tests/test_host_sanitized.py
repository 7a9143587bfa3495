"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
only): the host translation units of libwv_hip.so -- the .cube reader and
writers, the host Huffman coder and decoder, the split plan, the API's
argument checks -- rebuilt with host-side sanitizers
(wavelets_amd.build.build_sanitized), then the host test files run against
that build in a child process with the ASan runtime preloaded.  Any
sanitizer report fails the run (halt_on_error)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TESTS = ["tests/test_host_abi.py", "tests/test_host_decode.py", "tests/test_split.py"]


def test_host_code_clean_under_asan_and_ubsan(tmp_path):
    sys.path.insert(0, ROOT)
    from wavelets_amd import build as b
    if not os.path.exists(os.path.join(b.BUILD, "wv_dwt.hip.o")):
        pytest.skip("the library's objects are not built (python -m wavelets_amd.build)")
    lib = b.build_sanitized()
    rt = b.asan_runtime()
    assert rt and os.path.exists(rt), "no ASan runtime found"
    log = str(tmp_path / "san")
    env = dict(os.environ, LD_PRELOAD=rt, WV_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:log_path=" + log,
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:log_path=" + log)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu"] + HOST_TESTS, cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=900)
    reports = [open(os.path.join(tmp_path, f)).read() for f in os.listdir(tmp_path)
               if f.startswith("san")]
    assert not reports, "\n".join(reports)[:4000]
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert " passed" in r.stdout
