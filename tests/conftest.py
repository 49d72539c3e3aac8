"""Shared pytest configuration.

Markers:
  gpu  -- needs an MI355X (HIP device); run with ``-m gpu`` on a machine that has one.
  slow -- CPU tests that take tens of seconds (still part of the default run).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")
    config.addinivalue_line("markers", "slow: CPU test taking tens of seconds")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle  # test infrastructure: the FP64 CPU oracle

    oracle.lib()
    return oracle
