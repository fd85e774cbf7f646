"""The boundary's host code under AddressSanitizer (tests/asan): host
NumPy-style buffers staged through device scratch, odd sizes, unaligned
pointers, refusals, and eight threads calling at once.  The driver is a
separate process built with -Xarch_host -fsanitize=address (device code is
not instrumented); an ASan report aborts it with a non-zero status."""

import os
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
BIN = Path(__file__).resolve().parent / "asan" / "boundary_asan"


def test_boundary_under_asan(hip):
    if not BIN.exists():
        pytest.skip("tests/asan/boundary_asan not built (python -m numfast_amd.build --asan)")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0")
    print("running", BIN, flush=True)
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=240, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "boundary under ASan: ok" in r.stdout
