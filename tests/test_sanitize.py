"""ASan + UBSan run of the CPU oracle and the library's host-side C ABI paths
(tests/cpp/test_sanitize.cpp): no GPU needed. Built on first use, and again
whenever a source is newer than the driver; leak checking is on, with the
HIP runtime's own allocations suppressed."""
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _stale(exe):
    if not os.path.exists(exe):
        return True
    srcs = (glob.glob(os.path.join(ROOT, "moleculardynamics_amd", "csrc", "*"))
            + glob.glob(os.path.join(ROOT, "oracle", "*.cpp"))
            + glob.glob(os.path.join(ROOT, "oracle", "*.hpp"))
            + [os.path.join(ROOT, "tests", "cpp", "test_sanitize.cpp"),
               os.path.join(ROOT, "include", "moldyn_hip.h")])
    t = os.path.getmtime(exe)
    return any(os.path.getmtime(s) > t for s in srcs)


def test_sanitized_oracle_and_host_paths(tmp_path):
    from moleculardynamics_amd import build

    if not os.path.exists(build.HIPCC):
        pytest.skip("hipcc not available")
    exe = build.SAN_EXE
    if _stale(exe):
        exe = build.build_sanitized(verbose=False)
    supp = tmp_path / "lsan.supp"
    supp.write_text("leak:libamdhip64\nleak:libhsa-runtime64\nleak:librocprofiler-register\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               LSAN_OPTIONS=f"suppressions={supp}",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("sanitize ok")
