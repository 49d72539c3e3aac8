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


@pytest.fixture(scope="session")
def oracle_4096(oracle_mod):
    """The FP64 oracle on the N=4096 configuration (Re=1e5, K0=4, CFL=0.1, seed 0): 100
    steps with diagnostics every 100 (NEXTDT at it = 1 and 100), state at steps 0 and 100."""
    N, steps, update = 4096, 100, 100
    o = oracle_mod.Oracle(N, 1e5, 4.0, 0.1)
    o.init_config(0)
    om0 = o.om2()
    o.start(form=1)
    print(f"oracle N={N}: {steps} steps", flush=True)
    ro = o.run(steps, update, form=1)
    return {"om0": om0, "om": o.om2(), "rows": ro, "steps": steps, "update": update}


@pytest.fixture(scope="session")
def oracle_4096_uiuj(oracle_mod):
    """The same N=4096 run on the oracle in the reference's u_i u_j form (form 0)."""
    N, steps, update = 4096, 100, 100
    o = oracle_mod.Oracle(N, 1e5, 4.0, 0.1)
    o.init_config(0)
    o.start(form=0)
    print(f"oracle N={N} (u_i u_j form): {steps} steps", flush=True)
    ro = o.run(steps, update, form=0)
    return {"om": o.om2(), "rows": ro, "steps": steps, "update": update}
