"""Test configuration: the ``gpu`` marker, library fixtures and the oracle.

CPU tests (``-m "not gpu"``) check the oracle against the golden vectors,
the host-side model of the product library, and the C ABI exports.  GPU
tests (``-m gpu``) run the HIP path through the C ABI and compare it with
the oracle (and the reference library when it was built here).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libhotspot_ref.so")
PROD_LIB = os.path.join(ROOT, "hotspot_amd", "libhotspot_amd.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _gpu_available() -> bool:
    """Asks the HIP runtime for a device (no torch import: fresh GPU boxes
    page torch in slowly)."""
    import ctypes
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = ctypes.CDLL(name)
        except OSError:
            continue
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    return False


@pytest.fixture(scope="session")
def prod():
    """The product library; built here if missing."""
    from hotspot_amd import iface
    if not os.path.exists(PROD_LIB):
        from hotspot_amd import build
        build.build(verbose=False)
    lib = iface.HotSpotLib(PROD_LIB)
    lib.lib.hs_set_fatal_mode(0)
    return lib


@pytest.fixture(scope="session")
def ref():
    """The reference library compiled from its sources (absent on machines
    without the reference tree: tests that need it are skipped)."""
    from hotspot_amd import iface
    if not os.path.exists(REF_LIB):
        pytest.skip("reference library not built (oracle/_ref)")
    return iface.HotSpotLib(REF_LIB)


@pytest.fixture(scope="session")
def gpu():
    if not _gpu_available():
        pytest.skip("no GPU")
    return True
