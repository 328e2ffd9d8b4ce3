"""The divide-by-grid-step helper of the HIP kernels (lepic_qdiv.h) against
IEEE division, bit for bit, compiled for the host from the same header."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_qdiv_matches_ieee_division(tmp_path):
    exe = tmp_path / "qdiv_check"
    cc = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                         os.path.join(HERE, "qdiv_check.cpp"), "-o", str(exe)],
                        capture_output=True, text=True)
    if cc.returncode != 0:
        pytest.fail(cc.stderr)
    r = subprocess.run([str(exe), "20000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
