"""Shared test setup.

MKL's results depend on its thread count; the bit-exact comparisons against
the reference's fixtures are made with one MKL thread, as the fixtures were
produced.  This must be set before MKL is loaded.
"""
import os
import sys

os.environ.setdefault("MKL_NUM_THREADS", "1")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "gcge-1.0_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: long CPU test")
