"""The C++ host API (include/segmentctscans.hpp) drives the reference's
call sequence end to end on the device."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

EXE = Path(__file__).resolve().parent / "cpp" / "_build" / "api_example"


def test_cpp_api_example():
    assert EXE.exists(), "build first (__graft_entry__.build())"
    p = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    assert "api_example ok" in p.stdout
