"""HBM bandwidth probe: times the device copy kernels of libmfd_hbm_probe.so.

    python -m maxwellfdtd_amd.hbm_probe [--mib 1024] [--iters 5]

Prints the copy size and the achieved read+write bandwidth.  A copy of a known size is
also the calibration run for hardware byte counters (FETCH_SIZE / WRITE_SIZE): the
counters' reading for it, against its 2 x size bytes, gives the counters' scale.
"""
import argparse
import ctypes as C

from . import _lib


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=5)
    ap.add_argument("--plain-only", action="store_true",
                    help="only the plain copy (the byte-counter calibration run)")
    a = ap.parse_args()
    lib = _lib.probe()
    nbytes = a.mib << 20
    ms = C.c_double()
    lib.check(lib.probe_copy(nbytes, a.iters, C.byref(ms)), "probe_copy")
    print(f"copy {nbytes} bytes: read {nbytes} + write {nbytes} per launch, "
          f"{ms.value * 1e3:.1f} us, {2 * nbytes / ms.value / 1e6:.0f} GB/s", flush=True)
    if not a.plain_only:
        lib.check(lib.probe_copy_fast(nbytes, a.iters, C.byref(ms)), "probe_copy_fast")
        print(f"fast copy (4 x float4 per lane, nontemporal) {nbytes} bytes: "
              f"{ms.value * 1e3:.1f} us, {2 * nbytes / ms.value / 1e6:.0f} GB/s", flush=True)


if __name__ == "__main__":
    main()
