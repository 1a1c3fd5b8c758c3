"""Shared pytest setup: the `gpu` marker and import paths.

`-m "not gpu"` runs on any host (oracle pins, host logic, library exports);
`-m gpu` needs a HIP device and exercises the native library through its C ABI.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "edge-to_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")
