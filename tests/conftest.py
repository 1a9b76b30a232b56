"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
import sys
import threading
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

# The reference's Fortran routines use automatic arrays sized by NMAX; give
# threads that call them a large stack.
threading.stack_size(512 << 20)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.bindings import Oracle, ORC_PATH
    if not ORC_PATH.exists():
        from oracle.bindings import build
        build(reference=False)
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle.bindings import Reference, REF_PATH
    if not REF_PATH.exists():
        pytest.skip("reference build oracle/_ref not present")
    return Reference()


@pytest.fixture(scope="session")
def hip():
    from mercuda_amd.hip import MercuryHip
    lib = MercuryHip()
    if not lib.available():
        raise RuntimeError("GPU test run without a HIP device")
    return lib


def run_big_stack(fn, *args, **kw):
    """Run fn in a thread with a large stack (reference Fortran routines)."""
    out = {}

    def target():
        try:
            out["r"] = fn(*args, **kw)
        except BaseException as e:  # re-raised in the caller
            out["e"] = e

    t = threading.Thread(target=target)
    t.start()
    t.join()
    if "e" in out:
        raise out["e"]
    return out["r"]
