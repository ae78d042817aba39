"""The C++17 drop-in classes (dvmpp_amd/host) through the C++ test program
that mirrors the reference's GoogleTest suites (tests/cpp/test_dropin.cpp)."""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "tests", "cpp", "test_dropin.bin")


def test_cpp_dropin_suite():
    assert os.path.exists(BIN), "tests/cpp/test_dropin.bin not built (make -C dvmpp_amd/host)"
    r = subprocess.run([BIN, GOLDEN], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "0 failed" in r.stdout
