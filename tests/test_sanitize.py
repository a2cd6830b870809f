"""AddressSanitizer + UndefinedBehaviorSanitizer runs on the CPU: the oracle
(every entry point on small scenes, incl. depth-limit leaves and equal codes)
and the library's host code (scene generators, OBJ / MTL / .txt loaders with
malformed files, the SAH BVH host build, PNG / PPM writing, the render server
in dry-run mode with eviction, rejection and the batch window).  The builds
are tests/sanitize/Makefile (gcc for the oracle; hipcc with the sanitizers on
the host side only, linked by clang); no GPU is used."""
import os
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-C", HERE, "-j8"], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]


def _run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=600, env=ENV)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    return out


def test_oracle_under_sanitizers(built):
    assert "oracle sanitizer run ok" in _run([os.path.join(HERE, "_build", "oracle_driver")])


def test_host_code_under_sanitizers(built, tmp_path):
    assert "host sanitizer run ok" in _run([os.path.join(HERE, "_build", "host_driver"), str(tmp_path)])
