import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def kats():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")) as f:
        return json.load(f)


def _ensure_built():
    """Build the oracle (gcc) and, if absent, the gfx950 library (hipcc
    cross-compiles without a GPU) so that a fresh checkout can run the CPU
    suite."""
    import importlib.util
    from oracle import pyoracle
    if not os.path.exists(pyoracle._LIB_PATH):
        pyoracle.build()
    lib = os.path.join(ROOT, "spinoza_amd", "libspinoza_amd.so")
    if not os.path.exists(lib):
        spec = importlib.util.spec_from_file_location(
            "_spz_build", os.path.join(ROOT, "spinoza_amd", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()


_ensure_built()
