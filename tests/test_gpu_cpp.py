"""Runs the reference's known-answer tests written against the header-only
C++ API (tests/cpp/kat.cpp) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_cpp_kats(lslab):
    exe = os.path.join(HERE, "cpp", "kat")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp")], check=True)
    print("[case] cpp kat", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" ok") == 12
