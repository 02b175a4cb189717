"""The host side of the C ABI under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/sanitize/build.py builds the library with both sanitizers on its host
code and a C++ driver (tests/sanitize/host_check.cpp) that calls it.  On the
CPU the driver runs the argument checks and the no-device error paths; on a
GPU it runs every host-side driver (in-memory, z-chunked with threaded staged
copies, file-backed, the budget fallback, batched, device-resident, no
allocation, fields and EDT) and cross-checks their results bit for bit.  A
sanitizer report stops the driver with a nonzero status.
"""
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize"))
import build as san  # noqa: E402

REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")


def _check(mode):
    rc, out = san.run(mode)
    assert not any(r in out for r in REPORTS), out[-4000:]
    assert rc == 0, out[-4000:]
    assert "0 failure(s)" in out


def test_host_code_under_sanitizers_without_device():
    if san.stale():
        san.build()
    _check("cpu")


@pytest.mark.gpu
def test_host_code_under_sanitizers_on_device(gpu):
    if not os.path.exists(san.EXE):
        pytest.skip("sanitized build absent: run tests/sanitize/build.py before the GPU run")
    _check("gpu")
