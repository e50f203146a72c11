"""HBM calibration: a plain device copy of known size (lisa_copy_probe).

    python -m lisa_amd.copyprobe [--mib 1024] [--iters 10]

Prints the bytes moved per launch and the achieved GB/s.  Run under
`rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) to get the factor between
the counters and true bytes on this device.
"""
import argparse
import json

from lisa_amd import capi


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--iters", type=int, default=10)
    a = ap.parse_args()
    nbytes = a.mib << 20
    with capi.Context(0) as ctx:
        ms = ctx.copy_probe(nbytes, a.iters)
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "avg_ms": ms, "GBps_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}))


if __name__ == "__main__":
    main()
