"""Device-to-device copy of a known size through libcuspx.so, run under
``rocprofv3 --pmc FETCH_SIZE`` / ``--pmc WRITE_SIZE`` to scale those
counters to bytes (python3 -m cuspx_amd.calib)."""
import sys

import cuspx_amd as C


def main(mib=256, reps=3):
    n = mib << 20
    a, b = C.DeviceBuffer(n), C.DeviceBuffer(n)
    a.fill(1)
    for _ in range(reps):
        C._check(C.lib().cuspx_memcpy_dtod(b.ptr, a.ptr, n), "cuspx_memcpy_dtod")
    print("copied %d x %d bytes" % (reps, n), flush=True)


if __name__ == "__main__":
    main(*(int(x) for x in sys.argv[1:]))
