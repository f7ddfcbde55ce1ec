"""Test configuration: the `gpu` marker and import paths.

CPU tests (``-m "not gpu"``) cover the oracle against the reference build
and the golden fixtures, closed forms, host logic and the C-ABI exports.
GPU tests (``-m gpu``) are the parity tests proper and call the product
library through its C ABI.
"""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP library)")
