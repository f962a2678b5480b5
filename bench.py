#!/usr/bin/env python3
"""VV + critical-point throughput on one MI355X.

Workload: Kuhn 128^3 grid, 6 tets per cube (12,582,912 tets, 2,146,689
vertices), seeded vertex and tet permutations, +-0.2 cell jitter, float32
scalar = sum of 32 isotropic Gaussians + N(0, 1e-3).  One step = one pass of
the hot path: VV CSR (ascending rows) + classes[3V] recomputed from TV and
scalars already resident in HBM.

Prints one JSON line.  `roofline` prices the step against HBM with the
compulsory bytes 16T + 4V + 2(4(V+1) + 8E) + 12V (TV read, scalars read, VV
CSR written once and read once, classes written); the per-stage split of the
step comes from HIP events recorded between the pipeline's launches on the
stream the kernels run on.  `cpu_baseline` is the OpenMP CPU oracle on the
same mesh, with the thread count it used.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "tets/s VV+crit-pts @12.6M tets (% HBM roofline); batched meshes/s; p99 ms"
HBM_PEAK_GBS = 8000.0


def bytes_alg(T: int, V: int, E: int) -> int:
    return 16 * T + 4 * V + 2 * (4 * (V + 1) + 8 * E) + 12 * V


def load_traffic(workload: str):
    p = ROOT / "profiles" / f"traffic_{workload}.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return None


def cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def copy_calibration(torch, mib: int, reps: int) -> int:
    """Plain device copies of a known size: each dispatch reads and writes
    mib MiB, so a counter's reading per dispatch over mib MiB is its scale."""
    n = mib * (1 << 20) // 4
    a = torch.arange(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "dispatches": reps}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="Kuhn grid side")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the path is not sharded across devices "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import tetragpu_amd as t

    if args.copy_calibration:
        return copy_calibration(torch, args.copy_calibration, max(args.steps, 1))
    t.init(0)
    torch.cuda.init()
    n = args.n
    workload = f"kuhn{n}_gaussians"
    t0 = time.time()
    mesh = t.kuhn(n, seed=args.seed, scalar="gaussians")
    gen_s = time.time() - t0
    T, V = mesh.n_cells, mesh.n_points
    # a dedicated default-priority stream: the library launches on it and the
    # timing events are recorded on it
    stream = torch.cuda.Stream()
    d = t.Mesh(mesh.cells, V, mesh.scalars, stream=stream.cuda_stream)

    for _ in range(max(args.warmup, 1)):
        d.vv_critpoints()
    d.sync()
    nnz = d.shape("VV")["nnz"]
    E = nnz // 2

    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        d.vv_critpoints()
    stop.record(stream)
    torch.cuda.synchronize()
    d.sync()
    ms = start.elapsed_time(stop) / args.steps

    # per-stage split of one more step (events between launches)
    d.profile(True)
    d.vv_critpoints()
    stages = d.stage_times()
    d.profile(False)
    stage_total = sum(x for _, x in stages)
    dom_name, dom_ms = max(stages, key=lambda s: s[1]) if stages else ("", 0.0)

    B = bytes_alg(T, V, E)
    achieved = B / (ms * 1e-3) / 1e9
    traffic = load_traffic(workload)
    # The path's unit of work is one step (all launches); its compulsory bytes
    # are priced against HBM.  The dominant kernel is reported beside it with
    # its own event-timed duration and measured traffic.
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get("bytes_per_step") if traffic else None,
        "scope": "one step: VV CSR + classes from resident TV and scalars (all launches)",
        "bytes_alg": B,
        "bytes_alg_per_tet": round(B / T, 3),
        "dominant_kernel": {
            "name": dom_name, "ms": round(dom_ms, 4),
            "share": round(dom_ms / stage_total, 3) if stage_total else None,
            "traffic": traffic.get("dominant_kernel_bytes") if traffic else None,
            "traffic_gbs": (round(traffic["dominant_kernel_bytes"] / (dom_ms * 1e-3) / 1e9, 1)
                            if traffic and dom_ms > 0 else None),
        },
        "stages_ms": {k: round(v, 4) for k, v in stages},
    }

    cpu = None
    if not args.no_cpu:
        from oracle import binding as O

        threads = cpu_threads()
        _, _, secs = O.vv_cp_omp(mesh.cells, V, mesh.scalars, threads)
        cpu = {"value": round(T / secs, 1), "unit": "tets/s", "cores": threads,
               "kind": "port",
               "sample": f"full workload ({T} tets), one pass of the OpenMP oracle "
                         f"(VV + critical points), {secs:.2f} s"}
    d.close()

    line = {
        "metric": METRIC,
        "value": round(T / (ms * 1e-3), 1),
        "unit": "tets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": workload, "tets": T, "vertices": V, "edges": E,
                   "seed": args.seed, "scalar": "32 gaussians + N(0,1e-3), float32",
                   "generate_s": round(gen_s, 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
