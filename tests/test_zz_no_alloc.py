"""The allocation-free force (mercury_hip_force_noalloc) against the force of
the uploaded state (mercury_hip_force), bit for bit, on device buffers the
test allocates itself: two calls in a row, the second on a smaller system,
with 1PN and oblateness on (mfo_all's every term)."""
import ctypes as C

import numpy as np
import pytest

from mercuda_amd import systems

pytestmark = pytest.mark.gpu


class DeviceBuffers:
    """Plain hipMalloc'd buffers for the test's inputs and outputs, from the
    HIP runtime the library itself is linked against (symbols looked up through
    its handle: another runtime, such as the copy a torch wheel bundles, may be
    loaded in the same process)."""

    def __init__(self, runtime):
        self.hip = runtime
        self.ptrs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), C.c_size_t(max(arr.nbytes, 8))) == 0
        assert self.hip.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1) == 0
        self.ptrs.append(p)
        return p.value

    def get(self, ptr, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def soa(a, ld):
    """MERCURY-layout a[n, 3] -> component-major [3][ld]."""
    out = np.zeros((3, ld))
    out[:, :len(a)] = a.T
    return out


def run_case(hip, buf, sysd, pn, jcen, ld_extra):
    n, nb = len(sysd["m"]), sysd["nbig"]
    hip.configure(4)
    hip.upload_system(dict(sysd, jcen=np.asarray(jcen)), pn=pn)
    ref = hip.force()
    ld = n + ld_extra
    m = buf.put(sysd["m"])
    x = buf.put(soa(sysd["x"], ld))
    v = buf.put(soa(sysd["v"], ld))
    a = buf.put(np.full((3, ld), np.nan))
    hip.force_noalloc(n, nb, m, x, v, a, ld, pn=pn, jcen=jcen)
    got = buf.get(a, (3, ld))[:, :n].T
    print(f"case: n {n}, nbig {nb}, pn {pn}, jcen {list(jcen)}: "
          f"{np.count_nonzero(got != ref)} of {got.size} components differ", flush=True)
    assert np.array_equal(got, ref)


def test_noalloc_force_equals_force(hip):
    big = systems.config_small(seed=4, nparticles=3000)
    small = systems.config_small(seed=5, nparticles=700)
    jcen = (1.2e-9, -3.0e-13, 4.0e-17)
    buf = DeviceBuffers(hip.L)
    try:
        run_case(hip, buf, big, 1, jcen, 5)
        run_case(hip, buf, small, 1, jcen, 0)
        run_case(hip, buf, small, 0, (0.0, 0.0, 0.0), 3)
    finally:
        buf.free()
        hip.free()
