import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "smoq.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 device and libsmoq_amd.so")


@pytest.fixture(scope="session")
def ctx():
    import smoq_amd

    c = smoq_amd.default_context()
    print(f"[gpu] context on device {c.device}", flush=True)
    return c
