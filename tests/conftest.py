"""Test configuration: the `gpu` marker and shared fixtures.

CPU tests (-m "not gpu") exercise the oracle, the host-side builders of the
product library and its C ABI surface; GPU tests (-m gpu) are the parity tests
proper and call the HIP kernels through the C ABI.
"""
import os
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    import alien_amd._capi as capi
    import oracle
    if not capi.LIB_PATH.exists():
        from alien_amd import build
        build.build()
    if not oracle.LIB_PATH.exists():
        oracle.build()
    if not (ROOT / "tests" / "cxx" / "alien_api_check").exists():
        import subprocess
        subprocess.run(["make", "-s", "-C", str(ROOT / "tests" / "cxx")], check=True)


@pytest.fixture(scope="session", autouse=True)
def built_libraries():
    _ensure_built()
    yield


@pytest.fixture(scope="session")
def ctx():
    """One backend context (HIP stream + workspace) for the whole GPU session."""
    import alien_amd as aa
    c = aa.Context()
    yield c
    c.close()
