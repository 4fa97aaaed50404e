import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

GOLDEN = os.path.join(ROOT, "tests", "golden", "ref_vectors.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def po():
    import pyoracle

    pyoracle.orc()  # builds liboracle.so if needed
    return pyoracle


@pytest.fixture(scope="session")
def ref(po):
    r = po.ref()
    if r is None:
        pytest.skip("reference build oracle/_ref/libref.so not present (reference not mounted)")
    return r


@pytest.fixture(scope="session")
def gpu():
    import cufe_amd

    cufe_amd.lib()  # raises if the HIP library is missing: no fallback
    return cufe_amd


@pytest.fixture(scope="session")
def gpu_lib_optional():
    """The HIP library for host-only functions (CRT on the host), or None when
    it is not built; nothing here touches the device."""
    try:
        import cufe_amd

        cufe_amd.lib()
        return cufe_amd
    except Exception:
        return None
