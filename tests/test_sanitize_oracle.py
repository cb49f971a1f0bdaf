"""AddressSanitizer + UndefinedBehaviorSanitizer run of the CPU oracle: every
exported oracle routine is called over a small set of inputs, edge cases
included (very low T, large mu/T, both signs of xi, the (k0, k) = (0, 0)
cell, degenerate masses, a bad channel), from a driver compiled with the
oracle sources under -fsanitize=address,undefined (oracle/Makefile target
`sanitize`).  Any report aborts the driver, and leaks fail it at exit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _have_sanitizers():
    cc = shutil.which("gcc")
    if not cc:
        return False
    for lib in ("libasan.so", "libubsan.so"):
        p = subprocess.run([cc, f"-print-file-name={lib}"], capture_output=True, text=True).stdout.strip()
        if not os.path.isabs(p) or not os.path.exists(p):
            return False
    return True


@pytest.mark.skipif(not _have_sanitizers(), reason="gcc with libasan/libubsan not available")
def test_oracle_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True,
                       text=True, timeout=600)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert r.returncode == 0, out[-4000:]
    assert "no report" in out and "ERROR: AddressSanitizer" not in out and "runtime error" not in out
