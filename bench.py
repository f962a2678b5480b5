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
same mesh (1 warm-up + median of 5, host CPU model and SMT state stated),
beside the serial form that follows the reference's own builders.
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


def kernel_digest() -> str:
    """Digest of the HIP kernel sources: a counter profile applies to the
    build it was measured on, and is reported only while this matches."""
    import hashlib

    h = hashlib.sha256()
    for p in sorted((ROOT / "tetragpu_amd" / "csrc" / "hip").iterdir()):
        if p.suffix in (".hip", ".h"):
            h.update(p.name.encode())
            h.update(p.read_bytes())
    return h.hexdigest()[:16]


def load_profile(kind: str, workload: str):
    """profiles/<kind>_<workload>.json if it was measured on these kernel
    sources, else None."""
    p = ROOT / "profiles" / f"{kind}_{workload}.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
    except (OSError, ValueError):
        return None
    return d if d.get("kernel_digest") == kernel_digest() else None


def load_valu(workload: str):
    return load_profile("valu", workload)


def valu_busy(workload: str, dom_ms: float):
    v = load_valu(workload)
    if not v or dom_ms <= 0:
        return None
    k = next((x for n, x in v["kernels"].items() if n.startswith("k_bkt_vv_cp")), None)
    if not k:
        return None
    return round(k["SQ_INSTS_VALU"] * 4 / (dom_ms * 1e-3 * v["clock_hz"] * v["simds"]), 3)


def load_traffic(workload: str):
    return load_profile("traffic", workload)


# ---- rocprofv3 summaries (profiles/) ---------------------------------------------
PROFILE_RUN = ("python3 $R/bench.py --steps 2 --warmup 1 --hot-path-only")


def short_kernel(name: str) -> str:
    import re

    m = re.search(r"(k_\w+)(<[^>]*>)?", name)
    return (m.group(1) + (m.group(2) or "")) if m else name[:48]


def read_counters(path: Path) -> dict:
    """{kernel: {counter: [value per dispatch]}} from run_counter_collection.csv."""
    import csv
    from collections import defaultdict

    out = defaultdict(lambda: defaultdict(dict))
    with open(path) as f:
        for r in csv.DictReader(f):
            out[short_kernel(r["Kernel_Name"])][r["Counter_Name"]][int(r["Dispatch_Id"])] = \
                float(r["Counter_Value"])
    return {k: {c: list(v.values()) for c, v in d.items()} for k, d in out.items()}


def trace_summary(d: Path) -> int:
    """Markdown summary of a rocprofv3 kernel trace in d, over the window after
    the longest idle gap of at least half a second (the last phase of the run,
    e.g. the serving loop after its meshes were generated on the host): how much of the window some
    kernel was running (union of the kernel intervals), the mean number
    running at once, and the kernels with the most summed time."""
    import csv
    import re

    files = sorted(d.glob("**/*kernel_trace.csv"))
    if not files:
        raise SystemExit(f"no kernel trace under {d}")
    rows = []
    with open(files[0]) as fh:
        for r in csv.DictReader(fh):
            m = re.search(r"(k_[A-Za-z0-9_]+|__amd_rocclr_[A-Za-z]+)", r["Kernel_Name"])
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                         m.group(1) if m else r["Kernel_Name"][:48]))
    rows.sort()
    # busy segments (unions of overlapping kernels) and the longest gap between them
    segs = []
    for s, e, _ in rows:
        if segs and s <= segs[-1][1]:
            segs[-1][1] = max(segs[-1][1], e)
        else:
            segs.append([s, e])
    gaps = [(segs[i + 1][0] - segs[i][1], i + 1) for i in range(len(segs) - 1)]
    gap, first = max(gaps) if gaps else (0, 0)
    if gap < 500_000_000:  # no idle phase of half a second: the whole trace
        gap, first = 0, 0
    t0 = segs[first][0]
    win = [x for x in rows if x[0] >= t0]
    busy = sum(e - s for s, e in segs[first:])
    span = segs[-1][1] - t0
    per = {}
    for s, e, k in win:
        n, tot = per.get(k, (0, 0))
        per[k] = (n + 1, tot + e - s)
    total = sum(e - s for s, e, _ in win)
    print(f"window after the longest idle gap ({gap / 1e9:.3f} s): {len(win)} kernels over "
          f"{span / 1e9:.3f} s; some kernel running {busy / 1e9:.3f} s ({busy / span:.1%} of "
          f"the window); summed kernel time {total / 1e9:.3f} s ({total / busy:.2f} kernels "
          f"at once while busy)\n")
    print("| kernel | calls | summed ms | share | mean us |\n|---|---:|---:|---:|---:|")
    for k, (n, tot) in sorted(per.items(), key=lambda x: -x[1][1])[:14]:
        print(f"| {k} | {n} | {tot / 1e6:.1f} | {tot / total:.1%} | {tot / n / 1e3:.1f} |")
    return 0


def summarize(d: Path, workload: str) -> int:
    """Write profiles/*_<workload>.{json,md} from a directory of rocprofv3
    outputs made by the commands that the written files list:
    bench_FETCH_SIZE/ bench_WRITE_SIZE/ copy_FETCH_SIZE/ copy_WRITE_SIZE/
    (traffic), ktrace/ (kernel trace), pmc_valu/ (instruction counts of the
    bucket kernel, which bench.py turns into its VALU-busy share)."""
    import csv

    prof = ROOT / "profiles"
    # the digest of the kernel sources the runs measured, from their JSON lines
    digs = set()
    for log_file in d.glob("*.log"):
        for line in log_file.read_text(errors="replace").splitlines():
            if line.startswith("{") and '"kernel_digest"' in line:
                digs.add(json.loads(line)["kernel_digest"])
    if len(digs) > 1:
        raise SystemExit(f"runs in {d} measured different kernel sources: {sorted(digs)}")
    dig = digs.pop() if digs else kernel_digest()
    runs = None

    def n_runs(c):
        return max(len(v) for k, x in c.items() if k.startswith("k_bkt_vv_cp") for v in x.values())

    if all((d / f"{w}_{c}" / "run_counter_collection.csv").exists()
           for w in ("bench", "copy") for c in ("FETCH_SIZE", "WRITE_SIZE")):
        scale, per = {}, {}
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            cp = read_counters(d / f"copy_{c}" / "run_counter_collection.csv")
            # the copies only (the run also fills the source once)
            vals = [v for k, x in cp.items() if "copyBuffer" in k for v in x.get(c, [])]
            # a 1 GiB copy reads and writes 1 GiB: 1048576 KiB per dispatch
            scale[c] = 1048576.0 / (sum(vals) / len(vals))
            bc = read_counters(d / f"bench_{c}" / "run_counter_collection.csv")
            runs = n_runs(bc)
            per[c] = {k: sum(x.get(c, [])) * 1024 * scale[c] / runs for k, x in bc.items()}
        kernels = sorted(set(per["FETCH_SIZE"]) | set(per["WRITE_SIZE"]),
                         key=lambda k: -(per["FETCH_SIZE"].get(k, 0) + per["WRITE_SIZE"].get(k, 0)))
        dom = next(k for k in kernels if k.startswith("k_bkt_vv_cp"))
        fetch = sum(per["FETCH_SIZE"].values())
        write = sum(per["WRITE_SIZE"].values())
        j = {"workload": workload, "kernel_digest": dig, "hot_path_runs": runs,
             "bytes_per_step": round(fetch + write), "fetch_bytes_per_step": round(fetch),
             "write_bytes_per_step": round(write),
             "fetch_scale": round(scale["FETCH_SIZE"], 4),
             "write_scale": round(scale["WRITE_SIZE"], 4), "dominant_kernel": "k_bkt_vv_cp",
             "dominant_kernel_bytes": round(per["FETCH_SIZE"].get(dom, 0)
                                            + per["WRITE_SIZE"].get(dom, 0)),
             "kernels": {k: {"fetch": round(per["FETCH_SIZE"].get(k, 0)),
                             "write": round(per["WRITE_SIZE"].get(k, 0))} for k in kernels}}
        (prof / f"traffic_{workload}.json").write_text(json.dumps(j, indent=1) + "\n")
        lines = [f"# HBM traffic per step, {workload} (VV + critical points), MI355X", "",
                 "Commands (repository root `$R`, after `cd /tmp && export TMPDIR=/tmp`), "
                 "then `python3 $R/bench.py --summarize <dir>`:", "", "```"]
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            lines += [f"timeout -s KILL 120 rocprofv3 --pmc {c} --output-format csv "
                      f"-d <dir>/bench_{c} -o run -- {PROFILE_RUN}",
                      f"timeout -s KILL 90 rocprofv3 --pmc {c} --output-format csv "
                      f"-d <dir>/copy_{c} -o run -- python3 $R/bench.py --copy-calibration 1024 "
                      f"--steps 3"]
        lines += ["```", "",
                  f"Calibration: a 1 GiB device copy measured by the same counters gives the "
                  f"factors FETCH_SIZE x{scale['FETCH_SIZE']:.3f}, WRITE_SIZE "
                  f"x{scale['WRITE_SIZE']:.3f}.  Per step = counter sum over the run / "
                  f"{runs} hot-path runs.  Kernel sources digest `{dig}`.", "",
                  "| kernel (one step) | fetch MB | write MB |", "|---|---:|---:|"]
        for k in kernels:
            lines.append(f"| {k} | {per['FETCH_SIZE'].get(k, 0) / 1e6:.1f} | "
                         f"{per['WRITE_SIZE'].get(k, 0) / 1e6:.1f} |")
        lines.append(f"| **total** | **{fetch / 1e6:.1f}** | **{write / 1e6:.1f}** |")
        (prof / f"traffic_{workload}.md").write_text("\n".join(lines) + "\n")
        log(f"traffic: {(fetch + write) / 1e9:.3f} GB per step")

    if (d / "ktrace" / "run_kernel_stats.csv").exists():
        with open(d / "ktrace" / "run_kernel_stats.csv") as f:
            rows = list(csv.DictReader(f))
        lines = [f"# Kernel trace, {workload} (VV + critical points), MI355X", "",
                 "Command (repository root `$R`, after `cd /tmp && export TMPDIR=/tmp`), "
                 "then `python3 $R/bench.py --summarize <dir>`:", "", "```",
                 "timeout -s KILL 200 rocprofv3 --kernel-trace --stats --output-format csv "
                 "-d <dir>/ktrace -o run -- python3 $R/bench.py --steps 10 --warmup 2 "
                 "--hot-path-only", "```", "",
                 f"Kernel sources digest `{dig}`.", "",
                 "| kernel | calls | avg us | min us | max us | share % |",
                 "|---|---:|---:|---:|---:|---:|"]
        for r in rows:
            lines.append(f"| {short_kernel(r['Name'])} | {r['Calls']} | "
                         f"{float(r['AverageNs']) / 1e3:.1f} | {float(r['MinNs']) / 1e3:.1f} | "
                         f"{float(r['MaxNs']) / 1e3:.1f} | {float(r['Percentage']):.2f} |")
        (prof / f"kernels_{workload}.md").write_text("\n".join(lines) + "\n")

    if (d / "pmc_valu" / "run_counter_collection.csv").exists():
        cc = read_counters(d / "pmc_valu" / "run_counter_collection.csv")
        ks = {k: {c: sum(v) / len(v) for c, v in x.items()} for k, x in cc.items()
              if k.startswith("k_bkt_vv_cp")}
        (prof / f"valu_{workload}.json").write_text(json.dumps(
            {"workload": workload, "kernel_digest": dig, "clock_hz": 2.4e9, "simds": 1024,
             "command": "timeout -s KILL 90 rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_SALU "
                        "SQ_INSTS_LDS SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES --output-format csv "
                        f"-d <dir>/pmc_valu -o run -- {PROFILE_RUN}",
             "kernels": ks}, indent=1) + "\n")
        log(f"valu: {list(ks)}")
    return 0


def cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def cpu_info() -> dict:
    """Host CPU of the run: model name, logical CPUs, threads per core (SMT)."""
    model, siblings, cores = None, None, None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "siblings" and siblings is None:
                siblings = int(v)
            elif k == "cpu cores" and cores is None:
                cores = int(v)
    except OSError:
        pass
    smt = None
    try:
        smt = Path("/sys/devices/system/cpu/smt/active").read_text().strip() == "1"
    except OSError:
        pass
    tpc = siblings // cores if siblings and cores else None
    return {"model": model, "logical_cpus": os.cpu_count(), "threads_per_core": tpc,
            "smt_active": smt if smt is not None else (tpc > 1 if tpc else None),
            "affinity_cpus": cpu_threads()}


def median_time(fn, reps: int = 5) -> tuple:
    """One warm-up call, then the median of `reps` timed calls (steady wall
    clock); returns (median seconds, last result)."""
    import statistics

    res = fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        res = fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), res


def cpu_baselines(t, mesh, threads: int) -> dict:
    """The CPU restatement of the path (oracle/oracle.c), timed on this host:
    cpu_omp, the OpenMP form, on the full workload, and cpu_faithful, the
    serial form that follows the reference's find-based builders
    (src/cpu_extraction.cpp:259-289) plus the critPoints rule.  cpu_omp: 1
    warm-up + median of 5.  cpu_faithful: one run on the full workload (about
    18 s on 16-core EPYC hosts, so the sample stays bounded), beside 1 warm-up
    + median of 5 on Kuhn 48^3 with the same field."""
    import numpy as np
    from oracle import binding as O

    T, V = mesh.n_cells, mesh.n_points
    omp_s, (vv, cls, _) = median_time(lambda: O.vv_cp_omp(mesh.cells, V, mesh.scalars, threads))
    sc64 = mesh.scalars.astype(np.float64)
    t0 = time.perf_counter()
    O.critpoints(O.vv(mesh.cells, V), sc64, V)
    ser_full = time.perf_counter() - t0
    small = t.kuhn(48, seed=1234, scalar="gaussians")
    sc_small = small.scalars.astype(np.float64)
    ser_s, _ = median_time(lambda: O.critpoints(O.vv(small.cells, small.n_points), sc_small,
                                                small.n_points))
    return {"omp": {"value": round(T / omp_s, 1), "unit": "tets/s", "cores": threads,
                    "seconds_median": round(omp_s, 3), "median_of": 5,
                    "sample": f"full workload ({T} tets)"},
            "faithful": {"value": round(T / ser_full, 1), "unit": "tets/s", "cores": 1,
                         "seconds": round(ser_full, 3), "runs": 1,
                         "sample": f"full workload ({T} tets), one run",
                         "sample_kuhn48": {"value": round(small.n_cells / ser_s, 1),
                                           "seconds_median": round(ser_s, 3), "median_of": 5,
                                           "tets": small.n_cells}},
            "cpu": cpu_info(), "_ref": (vv, cls)}


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


def copy_ceiling(torch, mib: int = 1024, reps: int = 5) -> float:
    """Measured device-to-device copy rate (GB/s, read + write bytes), the
    practical ceiling for streaming kernels on this GPU: median of reps."""
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    times = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
    times.sort()
    del a, b
    return round(2 * n * 4 / (times[len(times) // 2] * 1e-3) / 1e9, 1)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(xs, q):
    import numpy as np
    return round(float(np.percentile(np.asarray(xs), q)), 3) if len(xs) else None


def gen_many(t, specs, threads=16):
    """Generate meshes in parallel (the generator is native; ctypes drops the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda kw: t.kuhn(**kw), specs))


def bench_batch(t, torch, n_meshes: int, reps: int, seed: int = 7) -> dict:
    """4096 Kuhn meshes, side log-uniform in [6, 40], uniform scalars: one
    mesh per launch sequence vs the batch path (waves, hipGraph replays)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = np.floor(np.exp(rng.uniform(np.log(6), np.log(41), n_meshes))).astype(int)
    t0 = time.time()
    meshes = gen_many(t, [dict(n=int(n), seed=seed * 100003 + i, scalar="uniform")
                          for i, n in enumerate(sides)])
    gen_s = time.time() - t0
    total_cells = int(sum(m.n_cells for m in meshes))
    log(f"batch: {n_meshes} meshes, {total_cells} cells generated in {gen_s:.1f} s")
    stream = torch.cuda.Stream()
    # plain path: one resident mesh per request, one launch sequence each
    singles = [t.Mesh(m.cells, m.n_points, m.scalars, stream=stream.cuda_stream) for m in meshes]
    for d in singles:
        d.vv_critpoints()
    torch.cuda.synchronize()
    loop_ms = []
    for _ in range(reps):
        a = time.perf_counter()
        for d in singles:
            d.vv_critpoints()
        stream.synchronize()
        loop_ms.append((time.perf_counter() - a) * 1e3)
    loop_cls = np.concatenate([d.classes() for d in singles])
    log(f"batch: one mesh per launch {min(loop_ms):.1f} ms")
    for d in singles:
        d.close()
    # batch path
    b = t.Batch([(m.cells, m.n_points, m.scalars) for m in meshes], stream=stream.cuda_stream)
    b.use_graph(True)
    b.vv_critpoints()  # eager warm-up (allocations)
    b.vv_critpoints()  # capture + first replay
    b.sync()
    batch_ms = []
    for _ in range(reps):
        a = time.perf_counter()
        b.vv_critpoints()
        stream.synchronize()
        batch_ms.append((time.perf_counter() - a) * 1e3)
    same = bool(np.array_equal(b.classes(), loop_cls))
    log(f"batch: batched {min(batch_ms):.1f} ms, identical={same}")
    waves = b.counts()["waves"]
    b.close()
    lm, bm = min(loop_ms), min(batch_ms)
    return {"meshes": n_meshes, "cells": total_cells, "waves": waves, "generate_s": round(gen_s, 1),
            "loop_ms": round(lm, 3), "batch_ms": round(bm, 3),
            "loop_meshes_per_s": round(n_meshes / (lm * 1e-3), 1),
            "meshes_per_s": round(n_meshes / (bm * 1e-3), 1),
            "speedup_vs_loop": round(lm / bm, 2), "tets_per_s": round(total_cells / (bm * 1e-3), 1),
            "identical_to_loop": same, "graph": True}


def bench_serve(t, rates, n_requests: int, seed: int = 11, slots: int = 8) -> dict:
    """Poisson arrivals of VV + critical-point requests (Kuhn grids with 10 %
    splits, T log-uniform in [10k, 5M], Gaussian-mixture scalars) followed by
    a VT + TT query on the resident mesh (built on the device, kept there);
    p50/p99 latency per rate, throughput, and the latency's phases."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np
    from tetragpu_amd.serve import TT, VT, Server

    rng = np.random.default_rng(seed)
    pool_n = 24
    targets = np.exp(rng.uniform(np.log(1e4), np.log(5e6), pool_n))
    sides = np.maximum(6, np.round((targets / (6 * 1.3)) ** (1 / 3))).astype(int)
    pool = gen_many(t, [dict(n=int(n), seed=seed * 7919 + i, split_fraction=0.1,
                             scalar="gaussians") for i, n in enumerate(sides)], threads=8)
    log(f"serve: pool of {pool_n} meshes generated")
    out = {"pool_cells_min": int(min(m.n_cells for m in pool)),
           "pool_cells_max": int(max(m.n_cells for m in pool)), "requests_per_rate": n_requests,
           "slots": slots,
           "rates": []}
    for rate in rates:
        lat1, lat2, cells = [], [], []
        phases = {k: [] for k in ("queue_ms", "upload_ms", "compute_ms", "followup_ms",
                                  "client_ms")}
        lock = threading.Lock()
        with Server(slots=slots, hbm_budget_bytes=96 << 30, fuse_max_cells=1 << 19,
                    fuse_max_meshes=64) as srv, ThreadPoolExecutor(64) as ex:
            def one(k, submitted):
                m = pool[k]
                rid = srv.submit(m.cells, m.n_points, m.scalars)
                r1 = srv.wait(rid)
                got = time.perf_counter()
                f = srv.followup(rid, VT | TT, copy=False)
                done = time.perf_counter()
                srv.release(rid)
                with lock:
                    lat1.append(r1["total_ms"])
                    lat2.append((done - submitted) * 1e3)
                    cells.append(m.n_cells)
                    for key in ("queue_ms", "upload_ms", "compute_ms"):
                        phases[key].append(r1[key])
                    phases["followup_ms"].append(f["followup_ms"])
                    # Python submit/wait/return around the server's own clock
                    phases["client_ms"].append((got - submitted) * 1e3 - r1["total_ms"])
            picks = rng.integers(0, pool_n, n_requests)
            gaps = rng.exponential(1.0 / rate, n_requests)
            start = time.perf_counter()
            nxt = start
            futs = []
            for k, g in zip(picks, gaps):
                nxt += g
                d = nxt - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                futs.append(ex.submit(one, int(k), time.perf_counter()))
            for f in futs:
                f.result()
            wall = time.perf_counter() - start
            groups = srv.stats()["groups"]
        thr = n_requests / wall
        log(f"serve: rate {rate}: p50 {pct(lat1, 50)} ms p99 {pct(lat1, 99)} ms, "
            f"{thr:.1f} req/s")
        out["rates"].append({"rate": rate, "p50_ms": pct(lat1, 50), "p99_ms": pct(lat1, 99),
                             "p50_ms_with_followup": pct(lat2, 50),
                             "p99_ms_with_followup": pct(lat2, 99),
                             "throughput_req_s": round(thr, 1),
                             # offered load not kept up with: the queue grows
                             # during the run and p99 measures that growth
                             "saturated": bool(thr < 0.95 * rate),
                             "phase_p50_ms": {k: pct(v, 50) for k, v in phases.items()},
                             "phase_p99_ms": {k: pct(v, 99) for k, v in phases.items()},
                             "tets_per_s": round(sum(cells) / wall, 1), "launches": groups})
    return out


def bench_relations(t, torch, mesh) -> dict:
    """First-build time of each relation on a fresh device mesh (prerequisites
    included, e.g. edge ids for EE), and of all of them in one request.

    Each relation is priced against HBM with its compulsory bytes: TV read
    once (16 bytes per cell) and the relation written once (int32 values; a
    CSR relation adds int32 offsets, rows + 1).  Intermediate relations a
    build needs (edge ids for EE, VT for TT) are not counted, so the fraction
    is an upper bound on how close the build is to the HBM roofline."""
    names = ["VE", "TE", "EV", "ET", "VF", "TF", "FV", "FE", "FT", "EF", "VT", "TT", "FF",
             "EE", "VV"]
    T = mesh.n_cells
    out = {"per_relation_ms": {}, "bytes_alg": {}, "hbm_frac": {},
           "bytes_rule": "16T + 4 x values + 4 x (rows + 1) for CSR relations"}

    def timed(mask, name=None):
        stream = torch.cuda.Stream()
        with t.Mesh(mesh.cells, mesh.n_points, mesh.scalars, stream=stream.cuda_stream) as d:
            d.sync()
            torch.cuda.synchronize()
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            d.build(mask)
            if mask == t.ALL_FLAGS:
                d.critpoints()
            b.record(stream)
            torch.cuda.synchronize()
            d.sync()
            shape = d.shape(name) if name else None
            return a.elapsed_time(b), shape

    for name in names:
        ms, sh = timed(t.flags(name), name)
        B = 16 * T + 4 * sh["nnz"] * sh["width"] + (0 if sh["fixed"] else 4 * (sh["rows"] + 1))
        out["per_relation_ms"][name] = round(ms, 3)
        out["bytes_alg"][name] = int(B)
        out["hbm_frac"][name] = round(B / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4)
    ms, _ = timed(t.ALL_FLAGS)
    out["all_flags_and_critpoints_ms"] = round(ms, 3)
    log(f"relations: all flags + critical points {ms:.2f} ms; HBM fractions {out['hbm_frac']}")
    return out


def bench_small(t, torch, steps: int = 200, check: bool = True) -> dict:
    """Kuhn 24^3 (82,944 tets) with U[0,1) scalars and seeded permutations:
    the CPU-runnable parity case.  It fits in L2/MALL, so it is latency-bound;
    reported as tets/s per step (mean over `steps` back-to-back steps)."""
    import numpy as np

    m = t.kuhn(24, seed=1234, scalar="uniform")
    stream = torch.cuda.Stream()
    out = {"workload": "kuhn24_uniform", "tets": m.n_cells, "vertices": m.n_points}
    with t.Mesh(m.cells, m.n_points, m.scalars, stream=stream.cuda_stream) as d:
        for _ in range(5):
            d.vv_critpoints()
        d.sync()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(steps):
            d.vv_critpoints()
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        cls = d.classes()
        off, vals = d.relation("VV")
    out.update({"ms_per_step": round(ms, 4), "tets_per_s": round(m.n_cells / (ms * 1e-3), 1)})
    if check:
        from oracle import binding as O

        th = cpu_threads()
        omp_s, (VV, ref, _) = median_time(lambda: O.vv_cp_omp(m.cells, m.n_points, m.scalars, th))
        sc64 = m.scalars.astype(np.float64)
        ser_s, _ = median_time(lambda: O.critpoints(O.vv(m.cells, m.n_points), sc64, m.n_points))
        out["identical_to_oracle"] = bool(np.array_equal(VV.off, off)
                                          and np.array_equal(VV.val[:, 0], vals[:, 0])
                                          and np.array_equal(ref, cls))
        out["cpu_omp"] = {"tets_per_s": round(m.n_cells / omp_s, 1), "cores": th,
                          "seconds_median": round(omp_s, 4), "median_of": 5}
        out["cpu_faithful"] = {"tets_per_s": round(m.n_cells / ser_s, 1), "cores": 1,
                               "seconds_median": round(ser_s, 4), "median_of": 5}
    log(f"small: {out}")
    return out


def bench_irregular(t, torch, steps: int, seed: int = 5, check: bool = True) -> dict:
    """Kuhn 96^3 with 15 % of the tets stellar-split 1->4 at their centroid,
    plus 32 hub vertices each raised by 64 repeated splits around them, so
    that the largest VV degree exceeds 60 (seeded vertex/tet permutation, 8
    sinusoidal modes): VV + critical points per step, then TT and EE built
    once on the resident mesh.  With check, VV and classes are compared with
    the OpenMP oracle on the full mesh (bit-exact)."""
    import numpy as np

    m = t.kuhn(96, seed=seed, split_fraction=0.15, n_hubs=32, hub_splits=64, scalar="sines")
    T, V = m.n_cells, m.n_points
    log(f"irregular: {T} tets, {V} vertices")
    stream = torch.cuda.Stream()
    out = {"workload": "kuhn96_split15_hubs32x64_sines", "tets": T, "vertices": V}
    with t.Mesh(m.cells, V, m.scalars, stream=stream.cuda_stream) as d:
        for _ in range(2):
            d.vv_critpoints()
        d.sync()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(steps):
            d.vv_critpoints()
        b.record(stream)
        torch.cuda.synchronize()
        d.sync()
        ms = a.elapsed_time(b) / steps
        d.profile(True)
        d.vv_critpoints()
        out["stages_ms"] = {k: round(v, 4) for k, v in d.stage_times()}
        d.profile(False)
        off, vals = d.relation("VV")
        deg = np.diff(off)
        cls = d.classes()
        out.update({"ms_per_step": round(ms, 4), "tets_per_s": round(T / (ms * 1e-3), 1),
                    "vv_nnz": int(off[-1]), "max_vv_degree": int(deg.max()),
                    "vertices_over_32": int((deg > 32).sum()),
                    "vertices_over_64": int((deg > 64).sum())})
        for name in ("TT", "EE"):
            torch.cuda.synchronize()
            a.record(stream)
            d.build(t.flags(name))
            b.record(stream)
            torch.cuda.synchronize()
            d.sync()
            out[f"{name.lower()}_build_ms"] = round(a.elapsed_time(b), 3)
            out[f"{name.lower()}_nnz"] = d.shape(name)["nnz"]
    if check:
        from oracle import binding as O

        VV, ref, secs = O.vv_cp_omp(m.cells, V, m.scalars, cpu_threads())
        same_vv = bool(np.array_equal(VV.off, off) and np.array_equal(VV.val[:, 0], vals[:, 0]))
        out["identical_to_oracle"] = bool(same_vv and np.array_equal(ref, cls))
        out["oracle_s"] = round(secs, 2)
    log(f"irregular: {out.get('ms_per_step')} ms per step, identical {out.get('identical_to_oracle')}")
    return out


def bench_translation(seed: int) -> dict:
    """Device->host transfer and host translation of the drop-in VV builder
    (make_VV_GPU: CSR copied back, then one std::vector per vertex), timed by
    the builder's own timers in the command-line driver, beside the builder's
    whole time (upload of TV excluded, build included).  The reference reports
    the same three intervals for its VV builder (README.md:115-123)."""
    import re
    import subprocess

    exe = Path(__file__).resolve().parent / "tetragpu_amd" / "lib" / "tetra_main"
    out = []
    for n in (24, 128):
        cmd = [str(exe), "-i", f"kuhn:{n}:{seed}:gaussians", "--no_validate", "-d", "1",
               "--build_VV", "-e", "/dev/null"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"{' '.join(cmd)} failed: {r.stderr[-500:]}")

        def grab(pat):
            m = re.search(pat, r.stdout)
            return float(m.group(1)) if m else None
        row = {"workload": f"kuhn{n}_gaussians", "tets": 6 * n ** 3,
               "vv_builder_s": grab(r"\[time\] VV \[GPU\]: ([0-9.eE+-]+) s"),
               "transfer_s": grab(r"Timer\[VV_GPU\] GPU Device->Host transfer: ([0-9.]+)"),
               "translation_s": grab(r"Timer\[VV_GPU\] GPU Device->Host translation: ([0-9.]+)")}
        out.append(row)
        log(f"translation: {row}")
    return {"builder": "make_VV_GPU via tetra_main", "meshes": out}


def bench_split(t, torch, mesh, whole, seed: int, div: int = 4) -> dict:
    """The automatic split (tg_critpoints_split) of the benchmark mesh under
    a device budget of the resident arrays plus 1/div of the whole mesh's
    footprint: vertex-id ranges, each classified on the cells that touch the
    range or its neighbours, must equal the whole-mesh run.  Run on the
    benchmark mesh (random vertex order: wide halos) and on the same grid
    with vertices and cells in grid order (the order of most mesh files)."""
    import numpy as np

    out = []
    n = round((mesh.n_cells / 6) ** (1 / 3))
    ordered = t.kuhn(n, seed=seed, permute_vertices=False, permute_cells=False,
                     scalar="gaussians")
    with t.Mesh(ordered.cells, ordered.n_points, ordered.scalars) as d:
        d.vv_critpoints()
        whole_ordered = d.classes()
    for name, m, ref in (("random_order", mesh, whole), ("grid_order", ordered, whole_ordered)):
        T, V = m.n_cells, m.n_points
        resident = 16 * T + 4 * V + 12 * V + 4 * (T + 1) + 4 * (V + 1) + V + (1 << 20)
        budget = resident + t.vv_footprint(T, V) // div
        sc = m.scalars.astype(np.float32)
        for _ in range(2):
            cls, st = t.split_classes(m.cells, V, sc, budget_bytes=budget)
        row = {"order": name, "budget_bytes": budget, "footprint_bytes": st["footprint_bytes"],
               "parts": st["parts"], "largest_part_cells": st["max_cells"],
               "cells_over_all_parts": st["cells_total"],
               "total_s": round(st["total_s"], 3), "localize_s": round(st["localize_s"], 3),
               "classify_s": round(st["classify_s"], 3),
               "identical_to_whole_mesh": bool(np.array_equal(cls, ref))}
        out.append(row)
        log(f"split: {row}")
    return {"rule": f"budget = resident arrays + footprint/{div}", "runs": out}


def bench_partitioned(t, mesh, whole, n_parts: int = 8) -> dict:
    """The partition loop (src/alg/critPoints.cu:654-1003) on the benchmark
    mesh cut into n_parts slabs of equal z extent.  Each slab is classified on
    the cells that touch it (a one-cell halo), one slab at a time on the GPU,
    so only the largest slab's mesh is resident.  Kuhn grids are flag
    complexes, so the merged classes must equal the whole-mesh run.  Times are
    host wall clock: localisation on the device (cell and vertex flags,
    scans, local ids), then per slab mesh creation + VV + classes + merge on
    the device (second of two runs)."""
    import numpy as np

    z = mesh.positions[:, 2].astype(np.float64)
    lo, hi = float(z.min()), float(z.max())
    parts = np.minimum(((z - lo) / max(hi - lo, 1e-30) * n_parts).astype(np.int32), n_parts - 1)
    sc = mesh.scalars.astype(np.float64)
    for _ in range(2):
        cls, st = t.partitioned_classes(mesh.cells, mesh.n_points, sc, parts)
    out = {"partitions": st["partitions"], "rule": "z slabs, cells touching the slab",
           "total_s": round(st["total_s"], 3), "localize_s": round(st["localize_s"], 3),
           "classify_s": round(st["classify_s"], 3), "largest_slab_cells": st["max_cells"],
           "cells_with_halos": st["cells_total"],
           "halo_overhead": round(st["cells_total"] / max(mesh.n_cells, 1) - 1.0, 4),
           "identical_to_whole_mesh": bool(np.array_equal(cls, whole))}
    log(f"partitioned: {out}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128, help="Kuhn grid side")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--variant", type=int, default=0,
                    help="kernel form of the hot path (0: the measured tuning table)")
    ap.add_argument("--bv-log", type=int, default=0,
                    help="bucketed forms: 2^k vertices per bucket (7 or 8; 0: tuning table)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--batch-meshes", type=int, default=4096,
                    help="meshes in the batching measurement (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=1000,
                    help="requests per arrival rate in the serving measurement (0: skip)")
    ap.add_argument("--rates", type=float, nargs="+", default=[100, 500, 1000, 2000, 5000])
    ap.add_argument("--serve-slots", type=int, default=8,
                    help="server slots (worker threads with their own streams)")
    ap.add_argument("--no-relations", action="store_true",
                    help="skip the per-relation build times on the large mesh")
    ap.add_argument("--no-irregular", action="store_true",
                    help="skip the irregular-degree mesh (Kuhn 96^3, 15 %% split)")
    ap.add_argument("--no-partitions", action="store_true",
                    help="skip the partition loop (8 z slabs of the benchmark mesh)")
    ap.add_argument("--profile-relations", action="store_true",
                    help="only build every relation flag + critical points on two fresh meshes "
                         "of the workload (for a kernel trace) and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (reference kernel for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--hot-path-only", action="store_true",
                    help="only the benchmark mesh's warm-up, timed and staged steps (for "
                         "counter profiles: every dispatch is the hot path's)")
    ap.add_argument("--trace-summary", type=Path, default=None, metavar="DIR",
                    help="print the busy share and the top kernels of a rocprofv3 kernel trace in DIR")
    ap.add_argument("--summarize", type=Path, default=None, metavar="DIR",
                    help="only write profiles/*.json|md from the rocprofv3 outputs in DIR "
                         "(no GPU)")
    args = ap.parse_args()
    if args.summarize:
        return summarize(args.summarize, f"kuhn{args.n}_gaussians")
    if args.trace_summary:
        return trace_summary(args.trace_summary)
    if args.hot_path_only:
        args.no_cpu = args.no_relations = args.no_irregular = args.no_partitions = True
        args.batch_meshes = args.serve_requests = 0
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
    from tetragpu_amd import tuning
    tuning.load()  # kernel form / shape per size class, measured on MI355X
    n = args.n
    workload = f"kuhn{n}_gaussians"
    t0 = time.time()
    mesh = t.kuhn(n, seed=args.seed, scalar="gaussians")
    if args.profile_relations:
        for _ in range(2):
            with t.Mesh(mesh.cells, mesh.n_points, mesh.scalars) as d:
                d.sync()
                t1 = time.perf_counter()
                d.build(t.ALL_FLAGS)
                d.vv_critpoints()
                d.sync()
                log(f"all flags + critical points {(time.perf_counter() - t1) * 1e3:.2f} ms")
        return 0
    gen_s = time.time() - t0
    T, V = mesh.n_cells, mesh.n_points
    # a dedicated default-priority stream: the library launches on it and the
    # timing events are recorded on it
    stream = torch.cuda.Stream()
    d = t.Mesh(mesh.cells, V, mesh.scalars, stream=stream.cuda_stream)
    if args.variant:
        d.set_config(variant=args.variant)
    if args.bv_log:
        d.set_bucket(args.bv_log, 8192)

    log(f"workload {workload}: {T} tets, {V} vertices")
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

    # K more steps with an event after each: the median step (the spread
    # behind the mean above)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    evs[0].record(stream)
    for i in range(args.steps):
        d.vv_critpoints()
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    per_step = sorted(evs[i].elapsed_time(evs[i + 1]) for i in range(args.steps))
    ms_median = per_step[len(per_step) // 2]

    # per-stage split of one more step (events between launches)
    d.profile(True)
    d.vv_critpoints()
    stages = d.stage_times()
    d.profile(False)
    stage_total = sum(x for _, x in stages)
    dom_name, dom_ms = max(stages, key=lambda s: s[1]) if stages else ("", 0.0)

    B = bytes_alg(T, V, E)
    B1 = 16 * T + 4 * V + 4 * (V + 1) + 8 * E + 12 * V
    achieved = B / (ms * 1e-3) / 1e9
    copy_gbs = copy_ceiling(torch)
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
        # the stricter single-pass figure (VV written once, not re-read)
        "bytes_alg_single_pass": B1,
        "frac_single_pass": round(B1 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        # a plain device copy measured in the same run
        "copy_ceiling_gbs": copy_gbs,
        "frac_of_copy_ceiling": round(achieved / copy_gbs, 4) if copy_gbs else None,
        "dominant_kernel": {
            "name": dom_name, "ms": round(dom_ms, 4),
            "share": round(dom_ms / stage_total, 3) if stage_total else None,
            "traffic": traffic.get("dominant_kernel_bytes") if traffic else None,
            "traffic_gbs": (round(traffic["dominant_kernel_bytes"] / (dom_ms * 1e-3) / 1e9, 1)
                            if traffic and dom_ms > 0 else None),
            # share of all SIMD cycles spent issuing VALU work (instruction
            # count from rocprofv3, duration live): the kernel's compute bound
            "valu_busy_frac": valu_busy(workload, dom_ms),
            # the step's compulsory bytes priced against this kernel alone:
            # the fraction the step could reach if every other launch took no
            # time
            "frac_if_alone": (round(B / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                              if dom_ms > 0 else None),
        },
        "stages_ms": {k: round(v, 4) for k, v in stages},
    }

    log(f"{ms:.4f} ms per step")
    cpu = None
    if not args.no_cpu:
        threads = cpu_threads()
        cb = cpu_baselines(t, mesh, threads)
        ref_vv, ref_cls = cb.pop("_ref")
        off, vals = d.relation("VV")
        same = bool(np.array_equal(off, ref_vv.off) and np.array_equal(vals[:, 0], ref_vv.val[:, 0])
                    and np.array_equal(d.classes(), ref_cls))
        omp = cb["omp"]
        cpu = {"value": omp["value"], "unit": "tets/s", "cores": threads, "kind": "port",
               "sample": f"full workload ({T} tets): the OpenMP oracle (VV + critical points), "
                         f"1 warm-up + median of 5, {omp['seconds_median']} s",
               "cpu_omp": omp, "cpu_faithful": cb["faithful"], "host": cb["cpu"],
               "gpu_identical_to_oracle": same}
        log(f"cpu baselines: omp {omp['value']:.3g} tets/s ({threads} threads), "
            f"faithful {cb['faithful']['value']:.3g} tets/s (1 thread)")
    d.close()

    if args.hot_path_only:
        print(json.dumps({"metric": METRIC, "value": round(T / (ms * 1e-3), 1), "unit": "tets/s",
                          "ms_per_step": round(ms, 4), "hot_path_only": True,
                          "hot_path_runs": max(args.warmup, 1) + 2 * args.steps + 1,
                          "kernel_digest": kernel_digest(), "roofline": roofline}))
        return 0
    # end to end from host buffers: upload TV + scalars, run, copy VV + classes back
    t0 = time.perf_counter()
    d2 = t.Mesh(mesh.cells, V, mesh.scalars, stream=stream.cuda_stream)
    d2.vv_critpoints()
    d2.relation("VV")
    whole_cls = d2.classes()
    e2e_ms = (time.perf_counter() - t0) * 1e3
    d2.close()

    partitioned = None if args.no_partitions else bench_partitioned(t, mesh, whole_cls)
    split = None if args.no_partitions else bench_split(t, torch, mesh, whole_cls, args.seed)

    relations = None if args.no_relations else bench_relations(t, torch, mesh)
    translation = None if args.no_relations else bench_translation(args.seed)
    small = bench_small(t, torch, check=not args.no_cpu)
    irregular = None if args.no_irregular else bench_irregular(t, torch, max(args.steps // 2, 1),
                                                                check=not args.no_cpu)
    batch = bench_batch(t, torch, args.batch_meshes, 3) if args.batch_meshes > 0 else None
    serve = (bench_serve(t, args.rates, args.serve_requests, slots=args.serve_slots)
             if args.serve_requests > 0 else None)

    line = {
        "metric": METRIC,
        "value": round(T / (ms * 1e-3), 1),
        "unit": "tets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "ms_per_step_median": round(ms_median, 4),
        "ms_per_step_min_max": [round(per_step[0], 4), round(per_step[-1], 4)],
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
        "end_to_end_ms": round(e2e_ms, 2),
        "end_to_end_note": "host cells/scalars in, VV CSR + classes out, incl. allocation (first use)",
        "relations": relations,
        "host_translation": translation,
        "small": small,
        "irregular": irregular,
        "partitioned": partitioned,
        "split": split,
        "batched": batch,
        "serving": serve,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
