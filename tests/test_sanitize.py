"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only):
the C ABI's no-device paths and the C++ API test's CPU cases against a
host-instrumented build of the library, and the CPU oracle
(tests/cpp/test_host_sanitize.cpp).  Any sanitizer report fails the run."""
import os
import subprocess

import pytest

from gpuequiv_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "tests", "cpp", "_build", "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _programs():
    progs = (os.path.join(OUT, "test_api_san"), os.path.join(OUT, "test_host_sanitize"))
    if not all(os.path.exists(p) for p in progs):
        progs = build.build_sanitized()
    return progs


def _device_present() -> bool:
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_device_present(), reason="the no-device paths need a machine without a GPU")
def test_host_code_under_sanitizers(tmp_path):
    api, host = _programs()
    r = subprocess.run([host], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    from tests.test_cpp_api import write_cases
    cases = tmp_path / "cases.txt"
    write_cases(str(cases), with_device_error=True)
    r = subprocess.run([api, "--cpu", str(cases)], capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
