"""Shared test setup: markers, import paths and the HIP context fixture."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device and the built HIP library")


@pytest.fixture(scope="session")
def hip():
    """The HIP backend context; a GPU test without a device is a failure, not a skip."""
    from echidna_amd.backend import HipContext

    ctx = HipContext.new(0)
    if ctx is None:
        pytest.fail("no gfx950 device visible: GPU tests need an MI355X (gfx950)")
    yield ctx
    ctx.close()
