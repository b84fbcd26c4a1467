"""Test configuration: the `gpu` marker and import paths.

CPU tests (``-m "not gpu"``) cover the oracle against golden vectors and
closed forms, the host logic, and that libkpm_hip.so loads and exports its
ABI.  GPU tests (``-m gpu``) are the parity tests of the HIP path and call
through the C ABI.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "kpm.jl_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
