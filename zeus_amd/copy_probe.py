"""Device-to-device copies of a known size, to calibrate the HBM traffic
counters of a profiler (FETCH_SIZE / WRITE_SIZE per dispatch) against the
bytes a plain copy must move, and to measure the copy bandwidth (bytes read
plus bytes written per second, best repetition).

    python -m zeus_amd.copy_probe [MiB] [repetitions]

Each repetition copies MiB mebibytes with zeus_memcpy_d2d (the runtime's
copy kernel): the kernel reads and writes exactly that many bytes, so
(bytes / 1024) / counter is the factor that turns the counter, in KiB per
dispatch, into bytes of HBM traffic.
"""
from __future__ import annotations

import ctypes as C
import sys
import time

from . import _capi as capi


def main(argv: list[str]) -> int:
    mib = int(argv[0]) if argv else 256
    reps = int(argv[1]) if len(argv) > 1 else 3
    lib = capi.lib()
    n = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    capi.check(lib.zeus_alloc(C.byref(src), n), "zeus_alloc")
    capi.check(lib.zeus_alloc(C.byref(dst), n), "zeus_alloc")
    try:
        best = float("inf")
        for _ in range(reps):
            capi.check(lib.zeus_synchronize(), "zeus_synchronize")
            t = time.perf_counter()
            capi.check(lib.zeus_memcpy_d2d(dst, src, n), "zeus_memcpy_d2d")
            capi.check(lib.zeus_synchronize(), "zeus_synchronize")
            best = min(best, time.perf_counter() - t)
    finally:
        lib.zeus_free(src)
        lib.zeus_free(dst)
    print(f"copied {mib} MiB x {reps} (device to device); best {best * 1e3:.3f} ms = "
          f"{2 * n / best / 1e9:.0f} GB/s read + write", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
