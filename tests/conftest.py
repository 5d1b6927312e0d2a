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


def assert_steps_close(got: dict, ref: dict, label: str = ""):
    """Parameters after the same Adam steps taken along two float orders of
    the same sums (graph vs plain, schedules, tilings, overlap, resume):
    per class normwise <= 1e-4, and at most 0.1 % of the elements outside
    rtol 1e-4 / atol 2e-6. Adam moves a parameter by ~lr sign(g) per step, so
    a gradient at roundoff level may take the other sign in one of the runs
    and move that element by up to lr per step."""
    import numpy as np
    for k in ref:
        a, b = np.asarray(got[k], np.float64), np.asarray(ref[k], np.float64)
        nrm = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
        off = float((np.abs(a - b) > 2e-6 + 1e-4 * np.abs(b)).mean())
        assert nrm <= 1e-4 and off <= 1e-3, f"{label} {k}: normwise {nrm:.2e}, elements off {off:.5f}"
