"""A plain device-to-device copy of known size: the calibration run for rocprofv3's HBM byte
counters.  Under `rocprofv3 --pmc FETCH_SIZE` (then WRITE_SIZE) each copy dispatch should read
(write) exactly `--mib` MiB; the ratio of that to the counter's reading is the factor applied to
the J/K kernels' counters.

    python3 -m byteqc_amd.perf.copy_probe --mib 1024 --reps 3
"""
from __future__ import annotations

import argparse
import sys


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split('\n\n')[0])
    ap.add_argument('--mib', type=int, default=1024)
    ap.add_argument('--reps', type=int, default=3)
    a = ap.parse_args(argv)
    import torch
    n = a.mib * (1 << 20) // 8
    x = torch.ones(n, dtype=torch.float64, device='cuda')
    y = torch.empty_like(x)
    for _ in range(a.reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f'copied {a.mib} MiB x {a.reps}', flush=True)
    return 0


if __name__ == '__main__':
    sys.exit(main())
