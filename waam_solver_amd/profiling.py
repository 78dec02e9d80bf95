"""Profiling helpers.

``python -m waam_solver_amd.profiling copy [--mib M] [--reps R]`` runs a plain device-to-device
copy of a known size R times. Hardware byte counters (rocprofv3 ``--pmc FETCH_SIZE`` /
``WRITE_SIZE``) read on this copy give the factor between what the counter reports and the
bytes actually moved, which scales the same counters read on the solver's kernels.

``summarize_counters(csv_path, kernel)`` averages one counter over the dispatches of a kernel in
a rocprofv3 ``counter_collection.csv``.
"""

from __future__ import annotations

import argparse
import csv
import sys


def copy_run(mib: int, reps: int) -> None:
    import torch

    n = mib * 1024 * 1024 // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[copy] {reps} copies of {mib} MiB ({n * 4} B read + {n * 4} B written each)")


def summarize_counters(csv_path: str, kernel: str) -> dict:
    """{counter: mean value per dispatch} over the dispatches whose name contains `kernel`."""
    sums: dict[str, float] = {}
    counts: dict[str, int] = {}
    with open(csv_path) as fh:
        for row in csv.DictReader(fh):
            if kernel not in row.get("Kernel_Name", ""):
                continue
            name = row["Counter_Name"]
            sums[name] = sums.get(name, 0.0) + float(row["Counter_Value"])
            counts[name] = counts.get(name, 0) + 1
    return {k: sums[k] / counts[k] for k in sums}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m waam_solver_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=256)
    c.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)
    if args.cmd == "copy":
        copy_run(args.mib, args.reps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
