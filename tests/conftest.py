import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X / gfx950)")


def load_cases(name):
    """Cases of a multi-volume fixture: list of (name, mask, kernel, reps, u, q)."""
    d = np.load(os.path.join(GOLDEN, name))
    out = []
    for n in d["names"]:
        n = str(n)
        out.append((n, d[f"{n}__mask"], d[f"{n}__kernel"], int(d[f"{n}__reps"]), d[f"{n}__u"],
                    d[f"{n}__q"]))
    return out


@pytest.fixture(scope="session")
def golden_small():
    return load_cases("small.npz")


@pytest.fixture(scope="session")
def golden_correlate():
    return load_cases("correlate.npz")


@pytest.fixture(scope="session")
def gpu():
    """The C-ABI library on a visible HIP device; skips when no device exists."""
    from maxibone_amd import capi
    capi.lib()  # raises if the library is not built: never a silent fallback
    if capi.device_count() < 1:
        pytest.skip("no HIP device visible")
    return capi


@pytest.fixture(scope="session")
def workload_mask():
    """The named workload's mask, built once per session (the 2048x2048x1536
    mask takes ~30 s and 6.4 GB)."""
    from maxibone_amd.synth import WORKLOADS
    cache = {}

    def get(name):
        if name not in cache:
            print(f"building the {name} mask", flush=True)
            cache[name] = WORKLOADS[name].mask()
        return cache[name]

    return get
