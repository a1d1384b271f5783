"""Sanitizer builds of the CPU restatement (oracle/pypo_oracle.cpp).

`make -C oracle sanitize` builds oracle/sanitize_main.cpp, which drives every
entry point of the restatement on small cases, twice: with AddressSanitizer
plus UndefinedBehaviorSanitizer, and with ThreadSanitizer (the targets and
rays are split over std::threads as Propagation.h:1356-1377 and
RayTrace.h:283-298 split them).  Any report ends the run with a non-zero
status (-fno-sanitize-recover).  sanitize_main.cpp also checks that one
thread and three threads give bit-identical results."""

import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.fixture(scope="module")
def built():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++ and make are needed for the sanitizer builds")
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(ORACLE, "_san")


@pytest.mark.parametrize("build", ["oracle_asan", "oracle_tsan"])
def test_restatement_clean_under_sanitizer(built, build):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(built, build)], capture_output=True, text=True, timeout=120,
                       env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "ALL OK" in r.stdout and "FAIL" not in r.stdout, out
    assert "runtime error" not in out and "WARNING: ThreadSanitizer" not in out, out
