"""Copy-kernel calibration run for HBM byte counters: launches the library's
16-byte copy kernel (lbm_copy_bandwidth) on 1 GiB buffers, so that each
dispatch reads exactly 1 GiB and writes exactly 1 GiB.  Profiling this run with
`rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` gives the counter reading of
a known byte count, the factor that converts the collide-stream kernels'
readings into HBM bytes.

    python lbmminiapp_amd/copy_calibration.py [bytes] [reps]
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from lbmminiapp_amd import capi  # noqa: E402


def main():
    nbytes = int(sys.argv[1]) if len(sys.argv) > 1 else 1 << 30
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 3
    gbs = capi.copy_bandwidth(nbytes, reps)
    print(f"copy {nbytes} bytes x {reps + 1} dispatches: {gbs:.1f} GB/s")


if __name__ == "__main__":
    main()
