"""Test configuration: the `gpu` marker and import paths.

`-m "not gpu"` runs the oracle checks, the host logic and the C-ABI export
checks on any machine; `-m gpu` runs the parity tests of the HIP engine
against the oracle on an MI355X.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
