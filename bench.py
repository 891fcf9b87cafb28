#!/usr/bin/env python3
"""Benchmark of the Faces exchange + DSS step on one MI355X.

Workload (N=1): the plain step on a 1x1x1 periodic rank grid of 60x60x60
elements of n=7 FP64 points (592.7 MB), integer-valued in [0,1023] from seed
42, every message addressed to self.  One step = one share, i.e. one pass of
the hot path over the field.  The field stays resident in HBM; the timed
region holds only the share launches: by default a HIP graph of ni = 10
shares (the reference's inner loop, gpu/main.cc:176-183) replayed steps/10
times, HIP events around every replay on the launch stream; --no-graph times
every share launch on its own.

value = algorithmic bytes per step / step time, with algorithmic bytes
B_step = 16 * mx*my*mz*n^3 (the field read once and written once).

Two further legs ride on the same line:
  "batched": 256 independent periodic grids (mx, my, mz uniform in [2,24], n
  from {4,6,8}, integer-valued [0,1023]) stepped by one work-list launch
  replayed from a HIP graph after the autotuner picked its shape, against the
  same grids stepped one launch per grid: grids/s, GB/s, and their ratio.
  "served": Poisson arrivals (mean 20k requests/s) of requests sized as the
  batched grids plus 5% large ones (48^3 elements, n=7), random FP64 [-1,1),
  fields resident in HBM, admission capped at 200 GB: p50/p99 latency and
  requests/s.

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
    "ragged": (7, 61, 60, 60, 0, "plain step 1x1x1 periodic, 61x60x60 elements (rows of 12 five-element "
                                "items and one of one element), n=7, integer-valued [0,1023] seed 42"),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--path", type=int, default=0, help="FACES_PATH_* (0 auto)")
    ap.add_argument("--rank-grid", default="1,1,1",
                    help="open boundary: LX,LY,LZ ranks, the plain leg stepping one of them (--rank), "
                         "its neighbours stood in for by self halos")
    ap.add_argument("--rank", type=int, default=0, help="the rank the plain leg steps (--rank-grid)")
    ap.add_argument("--boundary", default="periodic", choices=["periodic", "open"],
                    help="plain leg on one rank: periodic (default), or open (the reference's own mode)")
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
    ap.add_argument("--no-copy", action="store_true",
                    help="skip the copy and in-place stream references")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every share on its own instead of replaying a graph of --ni shares")
    ap.add_argument("--ni", type=int, default=10, help="shares per captured graph")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served leg")
    ap.add_argument("--grids", type=int, default=256, help="grids in the batched leg")
    ap.add_argument("--batch-tuning", default=None,
                    help="batched leg launch shape SEG,BLOCKS_PER_CU instead of the tuner's")
    ap.add_argument("--requests", type=int, default=2000, help="requests in the served leg")
    ap.add_argument("--streams", type=int, default=2, help="launch slots of the served leg")
    ap.add_argument("--serve-bpc", type=int, default=0,
                    help="served leg: blocks per CU of batched launches (-1: one per item; 0: default)")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--served-child", action="store_true", help=argparse.SUPPRESS)
    return ap.parse_args()


def splitmix64(seed, idx):
    """splitmix64 finaliser of seed ^ (idx * golden ratio): seeded shapes."""
    m = (1 << 64) - 1
    z = (seed ^ ((idx * 0x9E3779B97F4A7C15) & m)) & m
    z = (z + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def mixed_shapes(count, seed):
    """(n, mx, my, mz): mx, my, mz uniform in [2,24], n from {4,6,8}."""
    out = []
    for g in range(count):
        h = [splitmix64(seed, (g << 3) + k) for k in range(4)]
        out.append(((4, 6, 8)[h[3] % 3], 2 + h[0] % 23, 2 + h[1] % 23, 2 + h[2] % 23))
    return out


LARGE_SHAPE = (7, 48, 48, 48)  # the served workload's large requests (303.5 MB)


def served_workload(nreq, seed_shapes=5, seed_arrivals=20):
    """Arrival times (us) and shapes (n, mx, my, mz) of the served workload:
    Poisson arrivals with a mean of 50 us between requests (20k requests/s),
    sizes as the batched grids, 5% of them 48^3 n=7."""
    import random
    rng = random.Random(seed_arrivals)
    small = mixed_shapes(nreq, seed=seed_shapes)
    shapes, arrivals, t = [], [], 0.0
    for i in range(nreq):
        t += rng.expovariate(1.0 / 50.0)
        arrivals.append(t)
        shapes.append(LARGE_SHAPE if rng.random() < 0.05 else small[i])
    return arrivals, shapes


def event_ms(torch, fn, reps, stream):
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def batched_leg(torch, fa, ngrids, steps, warmup, tuning=None):
    from faces_amd.batch import Batch
    shapes = mixed_shapes(ngrids, seed=4)
    fields = [torch.empty(fa.field_doubles(*sh), dtype=torch.float64, device="cuda")
              for sh in shapes]

    def refill():
        for i, (d, sh) in enumerate(zip(fields, shapes)):
            fa.fill(d, *sh, kind=fa.FILL_INT, seed=42 + i)
        torch.cuda.synchronize()

    refill()
    b = Batch(shapes)
    b.set_fields(fields)
    stream = torch.cuda.current_stream()
    if tuning:  # fixed launch shape instead of the tuner: "SEG,BLOCKS_PER_CU"
        seg, bpc = (int(x) for x in tuning.split(","))
        b.set_tuning(256, seg, bpc)
        tuned = {"seg_len": seg, "blocks_per_cu": bpc}
    else:
        tuned = b.tune(reps=5, stream=stream)
    refill()
    for _ in range(warmup):
        b.share(stream=stream)
        b.share_per_grid(stream=stream)
    ms_b = event_ms(torch, lambda: b.share(stream=stream), steps, stream)
    ms_p = event_ms(torch, lambda: b.share_per_grid(stream=stream), max(5, steps // 4), stream)
    nbytes = sum(16.0 * n ** 3 * mx * my * mz for n, mx, my, mz in shapes)
    b.close()
    del fields
    torch.cuda.synchronize()
    return {"workload": f"{ngrids} periodic grids, mx,my,mz uniform [2,24], n from {{4,6,8}}, "
                        "integer-valued [0,1023]; one work-list launch per step (launch form in "
                        "'tuned': blocks_per_cu -1 is one item per block) replayed from a HIP graph",
            "grids": ngrids, "bytes_per_step": int(nbytes),
            "ms_per_step": round(ms_b, 5), "grids_per_s": round(ngrids / (ms_b * 1e-3), 1),
            "GBs": round(nbytes / (ms_b * 1e-3) / 1e9, 2),
            "per_grid_ms_per_step": round(ms_p, 5),
            "per_grid_grids_per_s": round(ngrids / (ms_p * 1e-3), 1),
            "per_grid_GBs": round(nbytes / (ms_p * 1e-3) / 1e9, 2),
            "speedup_vs_per_grid": round(ms_p / ms_b, 3),
            "tuned": {"seg_len": tuned["seg_len"], "blocks_per_cu": tuned["blocks_per_cu"]}}


def served_leg(torch, fa, nreq, args_nstreams=2, bpc=0):
    from faces_amd.batch import Server
    arrivals, shapes = served_workload(nreq)
    fields = []
    for i, sh in enumerate(shapes):
        d = torch.empty(fa.field_doubles(*sh), dtype=torch.float64, device="cuda")
        fa.fill(d, *sh, kind=fa.FILL_RAND, seed=7000 + i)
        fields.append(d)
    torch.cuda.synchronize()
    cap = 200e9
    srv = Server(footprint_cap=cap, max_batch=256, nstreams=args_nstreams)
    if bpc:
        srv.set_tuning(bpc)
    # warm the server's launch path on a few requests first, one of them
    # large (its plain-step plan is created when it is first submitted)
    warm = [0, 1, 2] + [i for i, sh in enumerate(shapes) if sh == LARGE_SHAPE][:1]
    srv.run_trace([10.0 * k for k in range(len(warm))], [shapes[i] for i in warm],
                  [fields[i] for i in warm])
    lat, st = srv.run_trace(arrivals, shapes, fields)
    # capacity: every request queued at once
    _, cap_st = srv.run_trace([0.0] * nreq, shapes, fields)
    srv.close()
    # launches overlapping on the device: up to 400 small requests, one per
    # launch, queued at once, through one launch slot and through two
    small_idx = [i for i, sh in enumerate(shapes) if sh != LARGE_SHAPE][:400]
    overlap = {}
    for ns in (1, 2):
        s1 = Server(footprint_cap=cap, max_batch=1, nstreams=ns)
        s1.run_trace([0.0] * 4, [shapes[i] for i in small_idx[:4]], [fields[i] for i in small_idx[:4]])
        _, st1 = s1.run_trace([0.0] * len(small_idx), [shapes[i] for i in small_idx],
                              [fields[i] for i in small_idx])
        overlap[ns] = st1["req_per_s"]
        s1.close()
    # the same requests one plain launch each, in arrival order
    plans = {}
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for sh, d in zip(shapes, fields):
        if sh not in plans:
            n, mx, my, mz = sh
            plans[sh] = fa.Faces(0, 1, 1, 1, mx, my, mz, n, boundary=fa.PERIODIC)
        plans[sh].share(d, stream=stream)
    torch.cuda.synchronize()
    per_req_s = time.perf_counter() - t0
    for f in plans.values():
        f.close()
    nbytes = sum(16.0 * n ** 3 * mx * my * mz for n, mx, my, mz in shapes)
    nlarge = sum(1 for sh in shapes if sh == LARGE_SHAPE)
    del fields
    torch.cuda.synchronize()
    return {"workload": f"{nreq} requests, Poisson arrivals mean 20k req/s, sizes as the batched "
                        f"grids plus 5% of 48^3 n=7 ({nlarge} drawn), random FP64 [-1,1), "
                        "fields resident in HBM, admission cap 200 GB",
            "requests": nreq, "offered_req_per_s": round(nreq / (arrivals[-1] * 1e-6), 1),
            "req_per_s": round(st["req_per_s"], 1), "p50_us": round(st["p50_us"], 1),
            "p99_us": round(st["p99_us"], 1), "max_us": round(st["max_us"], 1),
            "launches": int(st["launches"]), "footprint_cap_bytes": cap,
            "streams": args_nstreams, "blocks_per_cu": bpc or "default",
            "GBs": round(nbytes / st["wall_s"] / 1e9, 2),
            "capacity_req_per_s": round(cap_st["req_per_s"], 1),
            "per_request_launch_req_per_s": round(nreq / per_req_s, 1),
            "capacity_vs_per_request_launch": round(cap_st["req_per_s"] * per_req_s / nreq, 3),
            "one_request_per_launch_req_per_s": {"1_slot": round(overlap[1], 1), "2_slots": round(overlap[2], 1),
                                                 "requests": len(small_idx)}}


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


def inplace_stream_gbs(torch, fa, nbytes, reps=20):
    """Read+write bandwidth of the library's plain in-place streaming kernel
    over a buffer of the same size: the ceiling of an in-place step."""
    a = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    torch.cuda.synchronize()
    gbs = fa.stream_ceiling(a, reps=reps, stream=torch.cuda.current_stream())
    del a
    return gbs


def cpu_threads():
    """(threads, affinity cpus): one thread per CPU this process may run on,
    capped by OMP_NUM_THREADS where the host sets it (its CPU share)."""
    cpus = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(env) if env.isdigit() and int(env) > 0 else cpus
    return min(cpus, cap), cpus


def cpu_child(spec):
    """Runs in a child process with its OpenMP threads bound (see
    cpu_baseline): times the oracle's sweep and prints one JSON line."""
    import oracle
    n, mx, my, mz, kind, budget_s, periodic = json.loads(spec)
    u = oracle.new_field(n, mx, my, mz, kind, seed=42)
    t0 = time.perf_counter()
    oracle.share_sweep(u, n, mx, my, mz, periodic=periodic)
    first = time.perf_counter() - t0
    reps = max(1, min(50, int(budget_s / max(first, 1e-3))))
    t0 = time.perf_counter()
    for _ in range(reps):
        oracle.share_sweep(u, n, mx, my, mz, periodic=periodic)
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"dt": dt, "reps": reps, "threads": oracle.num_threads(),
                      "binding": oracle.omp_binding()}), flush=True)


def cpu_baseline(n, mx, my, mz, kind, budget_s, periodic=True):
    """The oracle's OpenMP sweep on the host cores: same workload, bounded.
    It runs in a child process whose OpenMP threads, one per core of this
    process's CPU share, are bound spread over the cores; binding them here
    would pin this process (and the server's threads) to one core."""
    import subprocess
    threads, cpus = cpu_threads()
    env = dict(os.environ, OMP_PROC_BIND="spread", OMP_PLACES="cores", OMP_NUM_THREADS=str(threads))
    spec = json.dumps([n, mx, my, mz, kind, budget_s, bool(periodic)])
    out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", spec], env=env,
                         capture_output=True, text=True, check=True, timeout=600)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    dt = r["dt"]
    b_step = 16.0 * n ** 3 * mx * my * mz
    return {"value": round(b_step / dt / 1e9, 3), "unit": "GB/s",
            "cores": r["threads"], "kind": "port",
            "nproc": os.cpu_count(), "affinity_cpus": cpus,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
            "binding": f"OMP_PROC_BIND=spread OMP_PLACES=cores (runtime: {r['binding']['proc_bind']}, "
                       f"{r['binding']['places']} places)",
            "ms_per_share": round(dt * 1e3, 3),
            "sample": f"{r['reps']} timed shares (+1 untimed) of the full {mx}x{my}x{mz} n={n} field "
                      f"by the oracle's OpenMP pass sweep; {dt * 1e3:.1f} ms/share"}


def load_traffic(name, plan_opts):
    """Measured HBM traffic of the step's kernel (profiles/traffic_<name>.json,
    from rocprofv3 FETCH_SIZE/WRITE_SIZE runs), with the kernel and commit it
    was measured on; None when the file is missing or was measured with other
    launch options than the plan's."""
    p = os.path.join(ROOT, "profiles", f"traffic_{name}.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    want = d.get("plan", {})
    if any(plan_opts.get(k) != v for k, v in want.items() if k in plan_opts):
        return None
    return {"bytes": d.get("hbm_bytes_per_launch"), "kernel": d.get("kernel"),
            "commit": d.get("measured_at_commit"), "source": d.get("source")}


def served_in_child(args):
    """The served leg in a child process of its own.  HIP shares a process's
    GPU_MAX_HW_QUEUES hardware queues (4 by default) among its streams; the
    streams the other legs leave behind (a graph-capture stream among them)
    would put the server's two launch slots on one queue and serialise them
    (p99 197-224 us against 161-167 us on queues of their own)."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--served-child", "--requests", str(args.requests),
           "--streams", str(args.streams), "--serve-bpc", str(args.serve_bpc)]
    out = subprocess.run(cmd, capture_output=True, text=True, check=True, timeout=600)
    return json.loads(out.stdout.strip().splitlines()[-1])


def main():
    args = parse()
    if args.cpu_child:
        cpu_child(args.cpu_child)
        return
    if args.served_child:
        import torch
        import faces_amd as fa
        print(json.dumps(served_leg(torch, fa, args.requests, args.streams, args.serve_bpc)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import faces_amd as fa

    n, mx, my, mz, kind, desc = CONFIGS[args.config]
    total = fa.field_doubles(n, mx, my, mz)
    b_step = 16.0 * n ** 3 * mx * my * mz
    # A field below 512 MB could stay in the 256 MB Infinity Cache between
    # steps: rotate enough fields that the working set exceeds it threefold.
    nbuf = 1 if 8 * total >= 512e6 else min(64, int(-(-768e6 // (8 * total))))
    us = [torch.empty(total, dtype=torch.float64, device="cuda") for _ in range(nbuf)]
    for x in us:
        fa.fill(x, n, mx, my, mz, kind=kind, seed=42)
    periodic = args.boundary == "periodic"
    lgrid = tuple(int(v) for v in args.rank_grid.split(","))
    if len(lgrid) != 3 or (periodic and lgrid != (1, 1, 1)):
        print("bench.py: --rank-grid takes LX,LY,LZ and needs --boundary open", file=sys.stderr)
        sys.exit(2)
    f = fa.Faces(args.rank, *lgrid, mx, my, mz, n, boundary=fa.PERIODIC if periodic else fa.OPEN)
    f.set_path(args.path)
    if args.block or args.seg >= 0 or args.xcd >= 0:
        f.set_tuning(args.block or 256, max(args.seg, 0), 1 if args.xcd < 0 else args.xcd)
    if args.variant or args.vblock:
        f.set_variant(args.variant, args.vblock)
    for o in args.opt:
        name, val = o.split("=")
        f.set_option(getattr(fa, "OPT_" + name.upper()), int(val))
    stream = torch.cuda.current_stream()
    plan_opts = {"flow": f.get_option(fa.OPT_FLOW), "xcd": f.get_option(fa.OPT_XCD),
                 "seg": f.get_option(fa.OPT_SEG)}

    for i in range(args.warmup):
        f.share(us[i % nbuf], stream=stream)
    torch.cuda.synchronize()

    if args.no_graph:
        # Events around every launch on the launch stream: per-launch durations.
        nrep, per_rep = args.steps, 1
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(nrep + 1)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        evs[0].record(stream)
        for i in range(nrep):
            f.share(us[i % nbuf], stream=stream)
            evs[i + 1].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    else:
        # A graph of ni shares (rounded up to whole turns of the rotated
        # fields), replayed; events around every replay.  When the step count
        # is not a multiple of that, one graph of all the steps, so that
        # exactly --steps shares are timed.
        per_rep = nbuf * max(1, -(-max(1, args.ni) // nbuf))
        if args.steps % per_rep:
            per_rep = max(1, args.steps)
        nrep = max(1, args.steps // per_rep)
        cap = torch.cuda.Stream()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=cap):
            for k in range(per_rep):
                f.share(us[k % nbuf], stream=cap)
        torch.cuda.synchronize()
        graph.replay()  # one untimed replay (the first is excluded, as in the reference)
        torch.cuda.synchronize()
        # replay() launches on the current stream: the events go there too
        rs = torch.cuda.current_stream()
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(nrep + 1)]
        t0 = time.perf_counter()
        evs[0].record(rs)
        for i in range(nrep):
            graph.replay()
            evs[i + 1].record(rs)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        args.steps = nrep * per_rep
    # measured traffic exists for the default path of one rank only
    traffic = (load_traffic(args.config if periodic else args.config + "_open", plan_opts)
               if args.path == 0 and lgrid == (1, 1, 1) else None)
    per = [evs[i].elapsed_time(evs[i + 1]) / per_rep for i in range(nrep)]
    ms = sum(per) / len(per)
    gbs = b_step / (ms * 1e-3) / 1e9
    f.close()
    del us
    torch.cuda.synchronize()

    copy_gbs = None if args.no_copy else copy_reference_gbs(torch, int(8 * total))
    stream_gbs = None if args.no_copy else inplace_stream_gbs(torch, fa, int(8 * total))
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, mx, my, mz, kind, args.cpu_seconds, periodic)

    batched = None if args.no_batch else batched_leg(torch, fa, args.grids, 50, 5, args.batch_tuning)
    served = None if args.no_serve else served_in_child(args)

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
        "config": {"workload": desc if periodic else
                   desc.replace("periodic", "open (the reference's own mode)", 1),
                   "n": n, "mx": mx, "my": my, "mz": mz,
                   "rank_grid": f"{'x'.join(map(str, lgrid))} {args.boundary}" + (f", rank {args.rank}" if lgrid != (1, 1, 1) else ""), "bytes_per_step": int(b_step),
                   "fields_rotated": nbuf,
                   "cache_resident": bool(nbuf * 8 * total < 512e6),
                   "path": "fused" if args.path in (0, 1) else f"path{args.path}",
                   "timing": "per launch" if args.no_graph
                   else (f"graph of {per_rep} shares replayed" if nbuf == 1 else
                         f"graph of {per_rep} shares over the {nbuf} fields in turn, replayed")},
        "steps_per_s": round(1e3 / ms, 1),
        "step_ms_min": round(per_sorted[0], 5),
        "step_ms_p50": round(per_sorted[len(per) // 2], 5),
        "wall_s": round(wall, 3),
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic["bytes"] if traffic else None,
                     "traffic_kernel": traffic["kernel"] if traffic else None,
                     "traffic_measured_at": traffic["commit"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None},
        "copy_reference_gbs": None if copy_gbs is None else round(copy_gbs, 1),
        "frac_of_copy": None if copy_gbs is None else round(gbs / copy_gbs, 4),
        "inplace_stream_gbs": None if stream_gbs is None else round(stream_gbs, 1),
        "frac_of_inplace_stream": None if stream_gbs is None else round(gbs / stream_gbs, 4),
        "cpu_baseline": cpu,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
