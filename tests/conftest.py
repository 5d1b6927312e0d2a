import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through the HIP C-ABI)")
    config.addinivalue_line("markers", "slow: longer CPU oracle checks")


@pytest.fixture(scope="session")
def gpu():
    """One GpuContext for the whole GPU session (runs through libgs_hip.so)."""
    import splattar_amd as sa
    ctx = sa.GpuContext(0)
    name, cus = ctx.device_info()
    print(f"\n[gpu] {name}, {cus} CUs", flush=True)
    yield ctx
    ctx.synchronize()
    ctx.close()


def announce(msg: str):
    """Name the case before each GPU step (a fault then still says what ran)."""
    print(f"[case] {msg}", flush=True)
