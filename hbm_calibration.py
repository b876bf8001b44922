"""Device copies of known size, for calibrating HBM traffic counters.

Under `rocprofv3 --pmc FETCH_SIZE` (and, in a second run, `WRITE_SIZE`) the
copy kernels below move exactly `--mib` MiB read and `--mib` MiB written per
launch; the counters' per-dispatch KiB over those sizes are the factors that
turn a kernel's counter readings into HBM bytes.

python hbm_calibration.py [--mib 1024] [--reps 3]   (from the repository root)

Measurement tooling beside bench.py; the library itself never imports torch.
"""
import argparse

import torch


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    n = args.mib * (1 << 20) // 4
    x = torch.rand(n, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(args.reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print("copied %d x %d MiB (read %d KiB, write %d KiB per copy)"
          % (args.reps, args.mib, args.mib * 1024, args.mib * 1024), flush=True)


if __name__ == "__main__":
    main()
