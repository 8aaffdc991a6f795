"""AddressSanitizer + UndefinedBehaviorSanitizer builds of host code, run on
the CPU (host compiler only, no GPU):

* the oracle: every entry point on small seeded grids with every array,
  source, stencil, density code, plane-wave kind and edge receivers
  (oracle/san_check.c, `make -C oracle san`);
* the solver library's host layout logic (seidart-backend_amd/csrc/
  host_layout.h): interior box, padded rows, the 3D elastic CPML-shell tile
  list and slots, material / coefficient dictionaries, padded host images
  and batch tile orders, with their invariants checked over seeded random
  grids (tests/host/layout_check.cpp, `make -C tests/host run`).

Either run aborts on the first sanitizer report (-fno-sanitize-recover) or
invariant failure.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(directory, target):
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, directory), target], capture_output=True, text=True,
                       env=env, timeout=600)
    out = r.stdout + r.stderr
    if r.returncode != 0 and ("cannot find -lasan" in out or "cannot find -lubsan" in out):
        pytest.skip("the host compiler has no sanitizer runtime")
    print(out[-3000:])
    assert r.returncode == 0, out[-3000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out
    return out


def test_oracle_under_sanitizers():
    out = _make("oracle", "san")
    assert "every oracle entry point ran clean" in out
    assert out.count("rc 0 unstable 0") == 33


def test_host_layout_under_sanitizers():
    out = _make("tests/host", "run")
    assert "layout_check: all checks passed" in out
