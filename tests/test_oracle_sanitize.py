"""The oracle and the device library's host-side ConfigFile parser under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only): `make -C oracle sanitize` builds oracle/sanitize_check.cpp with
-fsanitize=address,undefined and drives every oracle entry point on the parity geometries (sphere, z- and
x-cylinder, off-centre sphere, the reference's shipped case with its density filter, y- and z-normal inlets,
a periodic box) plus malformed ConfigFile text."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_oracle_and_config_parser_under_sanitizers():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "sanitize ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
