"""Copy-kernel calibration of HBM byte counters: copies 1 GiB between two device buffers with
the library's plain copy kernel (k_copy16) `repeats` times, so FETCH_SIZE / WRITE_SIZE read
under rocprofv3 can be scaled to known bytes (1 GiB read + 1 GiB written per dispatch).

    python -m markovkit_amd.calibrate [--bytes N] [--repeats R]
"""
import argparse
import ctypes as C
import time

from ._lib import check, lib


def copy_bandwidth(nbytes: int = 1 << 30, repeats: int = 5) -> float:
    """Best read+write bandwidth (GB/s, 2 * nbytes per copy) of the library's plain copy kernel
    between two device buffers of nbytes each."""
    L = lib()
    a, b = C.c_void_p(), C.c_void_p()
    check(L.mk_device_alloc(nbytes, C.byref(a)))
    check(L.mk_device_alloc(nbytes, C.byref(b)))
    try:
        check(L.mk_device_zero(a, nbytes))
        check(L.mk_device_copy(b, a, nbytes))  # warm-up
        check(L.mk_synchronize())
        best = float("inf")
        for _ in range(repeats):
            t = time.perf_counter()
            check(L.mk_device_copy(b, a, nbytes))
            check(L.mk_synchronize())
            best = min(best, time.perf_counter() - t)
    finally:
        check(L.mk_device_free(a))
        check(L.mk_device_free(b))
    return 2 * nbytes / best / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--repeats", type=int, default=5)
    args = ap.parse_args()
    L = lib()
    a, b = C.c_void_p(), C.c_void_p()
    check(L.mk_device_alloc(args.bytes, C.byref(a)))
    check(L.mk_device_alloc(args.bytes, C.byref(b)))
    check(L.mk_device_zero(a, args.bytes))
    check(L.mk_synchronize())
    for i in range(args.repeats):
        t = time.perf_counter()
        check(L.mk_device_copy(b, a, args.bytes))
        check(L.mk_synchronize())
        dt = time.perf_counter() - t
        print(f"copy {i}: {args.bytes} bytes in {dt * 1e3:.3f} ms "
              f"({2 * args.bytes / dt / 1e9:.0f} GB/s read+write)", flush=True)
    check(L.mk_device_free(a))
    check(L.mk_device_free(b))


if __name__ == "__main__":
    main()
