"""Test configuration: markers, paths and shared fixtures.

`gpu` tests need an MI355X and the built libswreax.so; everything else runs
on the CPU (oracle pins, host logic, C-ABI symbol checks).
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FFIELD = os.path.join(ROOT, "sw_reaxff_amd", "data", "ffield.reax")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: longer CPU test")


def pytest_runtest_logreport(report):
    """With SWR_INSTAFAIL=1 a failing test's traceback is printed when it
    fails, not only in the final summary (a long GPU run that is cut short
    keeps what it found)."""
    if report.failed and os.environ.get("SWR_INSTAFAIL"):
        print(f"\n[instafail] {report.nodeid} ({report.when})\n{report.longreprtext[-4000:]}", flush=True)


def _ensure_built(path, make_dir):
    if not os.path.exists(path):
        subprocess.run(["make", "-C", make_dir, "-j8"], check=True, stdout=subprocess.DEVNULL)


@pytest.fixture(scope="session")
def oracle_lib():
    _ensure_built(os.path.join(ROOT, "oracle", "liboracle_swreax.so"), os.path.join(ROOT, "oracle"))
    from oracle import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def ffield():
    return FFIELD


@pytest.fixture(scope="session")
def gpu_lib():
    from sw_reaxff_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        raise RuntimeError("libswreax.so is not built; run __graft_entry__.build()")
    return _lib
