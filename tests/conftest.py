import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "de-novo-antigen-binder_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def ctx():
    import dnab
    c = dnab.GpuContext.try_init(0)
    if c is None:
        pytest.fail("no usable gfx950 GPU for a test marked gpu")
    print(f"[gpu] {c.device_info()}", flush=True)
    yield c
    c.close()


@pytest.fixture(autouse=True)
def _name_the_case(request):
    # Print which test runs before any GPU work, so a crash can be attributed.
    if request.node.get_closest_marker("gpu"):
        print(f"\n[gpu-test] {request.node.nodeid}", flush=True)
    yield
