"""The C++ host layer (include/custinger_tc.hpp) driven like the reference's
drivers: tests/cct.cpp's launch-parameter sweep over the cuStinger and CSR
variants and tests/cct_upd.cpp's InsertionTest, each checked against the
oracle inside the driver (tests/cpp/tc_compat_driver.cpp)."""
from __future__ import annotations

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "tc_compat_driver")


def test_compat_driver():
    if not os.path.exists(_DRIVER):
        pytest.fail(f"{_DRIVER} is not built (run __graft_entry__.build())")
    r = subprocess.run([_DRIVER], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "compat driver ok" in r.stdout
