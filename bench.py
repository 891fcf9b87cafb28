#!/usr/bin/env python3
"""Benchmark of the Faces exchange + DSS step on one MI355X.

Workload (N=1): the plain step on a 1x1x1 periodic rank grid of 60x60x60
elements of n=7 FP64 points (592.7 MB), integer-valued in [0,1023] from seed
42, every message addressed to self.  One step = one share, i.e. one pass of
the hot path over the field.  The field stays resident in HBM; the timed
region holds only the share launches.

value = algorithmic bytes per step / step time, with algorithmic bytes
B_step = 16 * mx*my*mz*n^3 (the field read once and written once).

Prints ONE JSON line.  --gpus above 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "HBM GB/s & % roofline per exchange+DSS step; batched grids/s, p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth

CONFIGS = {
    # name: (n, mx, my, mz, fill kind, description)
    "c2": (7, 60, 60, 60, 0, "plain step 1x1x1 periodic, 60x60x60 elements, n=7, "
                            "integer-valued [0,1023] seed 42"),
    "c3": (8, 32, 32, 32, 1, "one rank of 6x6x6 periodic self-halo, 32x32x32 elements, n=8, "
                            "random [-1,1) seed 42"),
    "c1": (5, 4, 4, 4, 0, "plain step 1x1x1 periodic, 4x4x4 elements, n=5, integer seed 42"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--path", type=int, default=0, help="FACES_PATH_* (0 auto)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--seg", type=int, default=-1)
    ap.add_argument("--xcd", type=int, default=-1)
    ap.add_argument("--variant", type=int, default=0, help="FACES_VARIANT_* (0 auto)")
    ap.add_argument("--vblock", type=int, default=0,
                    help="LDS-kernel threads per block, or streaming-kernel blocks per CU")
    ap.add_argument("--opt", action="append", default=[],
                    help="plan option NAME=VALUE (faces_amd OPT_* without the prefix)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-copy", action="store_true", help="skip the copy-kernel reference")
    return ap.parse_args()


def copy_reference_gbs(torch, nbytes, reps=20):
    """Achieved bandwidth of a plain device copy of the same size (read+write)."""
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / reps
    del a, b
    return 2 * nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(n, mx, my, mz, kind, budget_s):
    """The oracle's OpenMP sweep on the host cores: same workload, bounded."""
    import numpy as np  # noqa: F401
    import oracle
    u = oracle.new_field(n, mx, my, mz, kind, seed=42)
    t0 = time.perf_counter()
    oracle.share_sweep(u, n, mx, my, mz, periodic=True)
    first = time.perf_counter() - t0
    reps = max(1, min(50, int(budget_s / max(first, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.share_sweep(u, n, mx, my, mz, periodic=True)
    dt = (time.perf_counter() - t0) / reps
    b_step = 16.0 * n ** 3 * mx * my * mz
    return {"value": round(b_step / dt / 1e9, 3), "unit": "GB/s",
            "cores": oracle.num_threads(), "kind": "port",
            "sample": f"{reps} timed shares (+1 untimed) of the full {mx}x{my}x{mz} n={n} field "
                      f"by the oracle's OpenMP pass sweep; {dt * 1e3:.1f} ms/share"}


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import faces_amd as fa

    n, mx, my, mz, kind, desc = CONFIGS[args.config]
    total = fa.field_doubles(n, mx, my, mz)
    b_step = 16.0 * n ** 3 * mx * my * mz
    u = torch.empty(total, dtype=torch.float64, device="cuda")
    fa.fill(u, n, mx, my, mz, kind=kind, seed=42)
    f = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
    f.set_path(args.path)
    if args.block or args.seg >= 0 or args.xcd >= 0:
        f.set_tuning(args.block or 256, max(args.seg, 0), 1 if args.xcd < 0 else args.xcd)
    if args.variant or args.vblock:
        f.set_variant(args.variant, args.vblock)
    for o in args.opt:
        name, val = o.split("=")
        f.set_option(getattr(fa, "OPT_" + name.upper()), int(val))
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        f.share(u, stream=stream)
    torch.cuda.synchronize()

    # Events around every launch on the launch stream: per-launch durations.
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(stream)
    for i in range(args.steps):
        f.share(u, stream=stream)
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps)]
    ms = sum(per) / len(per)
    gbs = b_step / (ms * 1e-3) / 1e9
    f.close()
    del u
    torch.cuda.synchronize()

    copy_gbs = None if args.no_copy else copy_reference_gbs(torch, int(8 * total))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, mx, my, mz, kind, args.cpu_seconds)

    per_sorted = sorted(per)
    rec = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "n": n, "mx": mx, "my": my, "mz": mz,
                   "rank_grid": "1x1x1 periodic", "bytes_per_step": int(b_step),
                   "path": "fused" if args.path in (0, 1) else f"path{args.path}"},
        "steps_per_s": round(1e3 / ms, 1),
        "step_ms_min": round(per_sorted[0], 5),
        "step_ms_p50": round(per_sorted[len(per) // 2], 5),
        "wall_s": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": load_traffic(args.config)},
        "copy_reference_gbs": None if copy_gbs is None else round(copy_gbs, 1),
        "frac_of_copy": None if copy_gbs is None else round(gbs / copy_gbs, 4),
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
