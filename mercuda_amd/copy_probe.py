"""A plain device-to-device copy of known size, for calibrating HBM byte
counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): the runtime's copy kernel
moves exactly `mib` MiB in and `mib` MiB out.

    python -m mercuda_amd.copy_probe 256
"""
from __future__ import annotations

import ctypes as C
import sys


def main(mib: int = 256) -> None:
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a = C.c_void_p(); b = C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise SystemExit("hipMalloc failed")
    hip.hipMemset(a, 1, C.c_size_t(n))
    hip.hipDeviceSynchronize()
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(n), 3) != 0:      # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a); hip.hipFree(b)
    print(f"copied {mib} MiB device-to-device 3 times")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 256)
