"""The C ABI without a GPU: the library loads, exports every function that
include/mercury_hip.h declares, the ctypes mirror covers exactly those, and
the host-side argument checks fail with an error return (never a crash, never
a CPU fallback).  Only calls that return before any device work are made."""
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mercury_hip.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mercury_hip_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from mercuda_amd import hip
    if not hip.LIB_PATH.exists():
        from mercuda_amd import build
        build.build()
    return hip.load_library()


def test_header_declares_the_reference_entry_points():
    """The 18 functions of mercury_cuda.h:13-33, renamed mercury_hip_*."""
    names = declared()
    ref18 = ["available", "configure", "reset", "limits", "export", "upload", "step", "download", "events",
             "force", "hybrid_begin", "hybrid_finish", "encounter_enter", "encounter_update", "encounter_exit",
             "free"]
    for n in ref18:
        assert f"mercury_hip_{n}" in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_mirror_matches_header():
    from mercuda_amd.hip import SIGNATURES
    assert sorted(SIGNATURES) == declared()


def test_host_side_errors_without_device(lib, capfd):
    """Unsupported algorithms, a step before any upload, bad tuning: nonzero
    returns with a message, and no device call is needed to reach them."""
    import ctypes as C
    for alg in (0, 5, 6, 7, 8, 11, -1):
        assert lib.mercury_hip_configure(alg) != 0    # not integrators of the reference's path
    for alg in (1, 2, 3, 4, 9):
        assert lib.mercury_hip_configure(alg) == 0    # MVS, BS, BS2, RADAU, TEST
    assert lib.mercury_hip_configure(10) == 0         # HYBRID
    assert lib.mercury_hip_encounter_update(None, None, None) != 0   # no encounter context
    h = C.c_double(4.0); hd = C.c_double(); nf = C.c_int64(); nr = C.c_int64()
    lib.mercury_hip_free()
    assert lib.mercury_hip_step(0.0, C.byref(h), C.byref(hd), 1e-12, C.byref(nf), C.byref(nr)) != 0
    x = np.zeros(3)
    assert lib.mercury_hip_download(x.ctypes.data_as(C.POINTER(C.c_double)),
                                    x.ctypes.data_as(C.POINTER(C.c_double)), 0) != 0
    assert lib.mercury_hip_set_tuning(0, 512, 0) != 0
    assert lib.mercury_hip_set_tuning(0, 96, 0) != 0
    assert lib.mercury_hip_set_tuning(0, 0, 9) != 0      # particles per lane 1..8
    assert lib.mercury_hip_set_tuning(0, 0, -1) != 0
    jc = np.zeros(3)
    jp = jc.ctypes.data_as(C.POINTER(C.c_double))
    dummy = C.c_void_p(1)
    assert lib.mercury_hip_force_noalloc(0, 1, 0, dummy, dummy, dummy, jp, dummy, 8) != 0
    assert lib.mercury_hip_force_noalloc(100, 65, 0, dummy, dummy, dummy, jp, dummy, 100) != 0
    assert lib.mercury_hip_force_noalloc(100, 9, 0, dummy, dummy, dummy, jp, dummy, 99) != 0
    err = capfd.readouterr().err
    assert "no state uploaded" in err and "block size" in err and "no encounter context" in err
    assert "scratch holds (64)" in err and "leading dimension" in err


def test_nongrav_upload_refusals_without_device(lib, capfd):
    """The non-gravitational model's limits, refused before any device call:
    an invalid flag; MVS with cometary or Yarkovsky parameters on a Big body
    (the MVS map takes them on particles only)."""
    import ctypes as C
    n, nbig = 4, 2
    m = np.array([2.959122082855911e-4, 1e-9, 0.0, 0.0])
    x = np.array([[0, 0, 0], [1.0, 0, 0], [2.0, 0, 0], [3.0, 0, 0]])
    v = np.array([[0, 0, 0], [0, 0.017, 0], [0, 0.012, 0], [0, 0.01, 0]])
    z3 = np.zeros(3); zn = np.zeros(n)
    dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))
    def upload(flag, ngf):
        return lib.mercury_hip_upload(n, nbig, 0, flag, dp(m), dp(x), dp(v), dp(ngf), dp(z3), dp(zn), dp(zn))
    on_particle = np.zeros((n, 5)); on_particle[2, 0] = 1e-8
    on_big = np.zeros((n, 5)); on_big[1, 4] = 1e-11
    assert lib.mercury_hip_configure(4) == 0
    assert upload(4, on_particle) != 0
    assert lib.mercury_hip_configure(1) == 0
    assert upload(3, on_big) != 0
    err = capfd.readouterr().err
    assert "invalid non-gravitational flag" in err
    assert "non-gravitational forces on Big bodies" in err


def test_no_device_means_not_available(lib):
    """On a machine without a GPU the library says so instead of computing."""
    if Path("/dev/kfd").exists():
        pytest.skip("a GPU driver is present")
    assert lib.mercury_hip_available() == 0
