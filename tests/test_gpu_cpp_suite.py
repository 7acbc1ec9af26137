"""Runs the C++ restatement of the reference's unit tests (tests/cpp), built
against the binder:: host interface, on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parent / "cpp" / "test_reference_suite"


def test_reference_suite_cpp():
    assert EXE.exists(), f"{EXE} not built (make -C tests/cpp)"
    r = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
