"""Race check of the oracle's parallel step (CPU): the oracle and a driver
(oracle/tsan_driver.c) built with ThreadSanitizer against LLVM's OpenMP
runtime, whose Archer tool makes the runtime's own synchronisation visible to
TSAN, and stepped with 4 threads: the per-particle loop, collisions, hits, the
triangle counters and the node charge scatter run concurrently.  The clean
build reports no race; a build with the charge scatter's atomic add planted
as a plain add must report one."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang"
ARCHER = "/opt/rocm/llvm/lib/libarcher.so"

pytestmark = pytest.mark.skipif(not (os.path.exists(CLANG) and os.path.exists(ARCHER)),
                                reason="LLVM clang with the Archer OpenMP tool not available")


@pytest.fixture(scope="module")
def drivers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "tsan"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(ROOT, "oracle", "_tsan")


def _run(exe):
    env = dict(os.environ, OMP_TOOL_LIBRARIES=ARCHER,
               TSAN_OPTIONS="halt_on_error=0 ignore_noninstrumented_modules=1 exitcode=66")
    return subprocess.run([exe, "4"], capture_output=True, text=True, timeout=600, env=env)


def test_parallel_step_has_no_race(drivers):
    r = _run(os.path.join(drivers, "driver"))
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[:4000]
    assert r.returncode == 0, r.stdout + r.stderr[:2000]
    f = r.stdout.split()
    assert f[0] == "steps" and int(f[5]) > 0  # particles settled, so the hit path ran


def test_planted_race_is_reported(drivers):
    r = _run(os.path.join(drivers, "driver_planted"))
    assert r.returncode == 66 and "WARNING: ThreadSanitizer: data race" in r.stderr
    assert "atomic_add_double" in r.stderr
