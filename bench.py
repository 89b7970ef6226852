#!/usr/bin/env python3
"""Benchmark: incremental delta-PageRank / SSSP after edge-update batches on one
MI355X (GTEPS, fraction of the HBM roofline, speedup vs full recompute on the
same GPU, and the CPU oracle timed on the host as a baseline).

A step = apply one seeded update batch (device CSR/CSC delta-merge + reseed or
SSSP repair) and run the incremental queries to their fixed points.  The graph is
generated on the device and the batches are uploaded to device memory before the
timed region, so the timed steps read their inputs from HBM.

Default workload (--workload web26, the metric's large configuration): directed
R-MAT scale 26 (67.1M vertices, a=.57 b=.19 c=.19, edge factor 16, seed 5,
self-loops/duplicates removed), u32 weights uniform in [1,255], supernodes from
leaf clustering (at most 1024 members, compress threshold 1), successive seeded 1%
batches (half deletes of existing edges, half inserts with uniform endpoints); a
step brings FP32/Kahan delta-PageRank (the SumInc query over the sketch, FP64
residual, T = 1e-7) and SSSP over the sketch (source 0) back to their fixed points.
The other workloads are the metric's smaller configurations (--workload pr16 |
sssp22 | pr22 | pr24 | pr26 | wpr22).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Δ-PR/SSSP GTEPS @L1≤1e-6 vs FP64 on 1 MI355X, %HBM peak, vs full recompute"
HBM_PEAK_GBS = 8000.0
DAMPING = float(np.float32(0.85))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


WORKLOADS = {
    "sssp22": dict(kind="sssp", scale=22, seed=2, wmin=1, wmax=64, frac=0.001, source=0, max_node=1024,
                   desc="incremental SSSP with supernodes, R-MAT s22 ef16 seed2, u32 w~U[1,64], source 0, "
                        "clusters <= 1024 members, 0.1% batches"),
    "pr24": dict(kind="pr", scale=24, seed=3, frac=0.01, T=1e-7, max_node=1024,
                 desc="incremental FP32/Kahan delta-PageRank, R-MAT s24 ef16 seed3, d=0.85f, 1% batches"),
    "pr22": dict(kind="pr", scale=22, seed=3, frac=0.01, T=1e-7, max_node=1024,
                 desc="incremental FP32/Kahan delta-PageRank, R-MAT s22 ef16 seed3, d=0.85f, 1% batches"),
    "pr26": dict(kind="pr", scale=26, seed=5, frac=0.01, T=1e-7, max_node=1024,
                 desc="incremental FP32/Kahan delta-PageRank, R-MAT s26 ef16 seed5, d=0.85f, 1% batches"),
    "pr16": dict(kind="prfull", scale=16, seed=1, eps=1e-9,
                 desc="plain delta-PageRank from scratch, R-MAT s16 ef16 seed1, d=0.85f, per-vertex threshold 1e-9 "
                      "(T = 1e-9 N), FP64, no compression, no updates (the reference's CPU case)"),
    "wpr22": dict(kind="wpr", scale=22, seed=4, seeds=1024, T=1e-9, fd_edges=256, h=1e-3,
                  desc="differentiable weighted PageRank, R-MAT s22 ef16 seed4, FP32 edge weights U(0,1] "
                       "row-normalized, personalization Dirichlet(1) over 1024 random seeds, loss sum c_i r_i "
                       "with c ~ N(0,1); a step = forward + adjoint (gradients w.r.t. weights and p)"),
    "web26": dict(kind="web", scale=26, seed=5, wmin=1, wmax=255, frac=0.01, T=1e-7, source=0, max_node=1024,
                  adjoint_seeds=1024,
                  desc="web-like incremental run, R-MAT s26 ef16 seed5 (a=.57 b=.19 c=.19), supernode compression "
                       "(leaf clusters <= 1024 members, compress threshold 1), successive seeded 1% batches (half "
                       "deletes of existing edges, half inserts); per batch FP32/Kahan delta-PageRank with an FP64 "
                       "residual (the SumInc query over the sketch, T = 1e-7) + SSSP over the sketch (u32 "
                       "w~U[1,255], source 0); weighted-PageRank forward + adjoint on the same graph"),
}


def pull_kw(args) -> dict:
    """The pull layout parameters of every context this script opens."""
    return dict(pull_band_kib=args.pull_band_kib, pull_tile_kib=args.pull_tile_kib, pull_bin_kib=args.pull_bin_kib,
                pull_bin_direct=args.pull_bin_direct)


def pull_mode(args, n) -> str:
    """Which pull the library's dense PageRank rounds run (its rules: the binned pull
    when a window is given, or the FP32 contributions exceed 8 MiB and it is not
    off; else the tiled pull when a window is given or they exceed 64 MiB; else the
    row classes)."""
    if args.pull_band_kib > 0:
        return "rows"
    if args.pull_bin_kib > 0 or (args.pull_bin_kib == 0 and n * 4 > (8 << 20)):
        return "binned"
    if args.pull_tile_kib > 0 or (args.pull_tile_kib == 0 and n * 4 > (64 << 20)):
        return "tiled"
    return "rows"


PULL_KERNELS = {"rows": "k_pr_pull", "tiled": "k_pr_pull_tiled", "binned": "k_pr_bin_gather+k_pr_bin_sum"}


def name_pull(args, n, kernel, roof, table):
    """The profiled pull launches are the tiled or binned kernels' when those run."""
    name = PULL_KERNELS[pull_mode(args, n)]
    if name == "k_pr_pull":
        return kernel
    if roof is not None and roof.get("kernel") == "k_pr_pull":
        roof["kernel"] = name
    if table and "k_pr_pull" in table:
        table[name] = table.pop("k_pr_pull")
    return name if kernel == "k_pr_pull" else kernel


def kernel_names():
    import suminc_amd._lib as L

    return {L.KERNEL_PR_PUSH: "k_pr_push", L.KERNEL_PR_GATHER: "k_pr_gather",
            L.KERNEL_SSSP_PUSH: "k_sssp_push+k_sssp_heavy", L.KERNEL_SSSP_SKETCH: "k_ss_push",
            L.KERNEL_PR_PULL: "k_pr_pull", L.KERNEL_BATCH_COPY: "k_copy_old"}


KERNEL_NAMES = {}


def roofline(ctx, kernel_ids, workload):
    import suminc_amd._lib as L

    best = None
    for k in kernel_ids:
        p = ctx.profile(k)
        if p["launches"] and (best is None or p["ms"] > best[1]["ms"]):
            best = (k, p)
    if best is None:
        return None, None
    k, p = best
    achieved = p["alg_bytes"] / (p["ms"] * 1e-3) / 1e9
    name = KERNEL_NAMES[k]
    alg = p["alg_bytes"] / p["launches"]
    # web26 measures the traffic live (run_web); the other workloads read the
    # committed counter summary of the same command
    traffic = None if workload == "web26" else measured_traffic(name, workload)
    return name, {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None if traffic is None else round(traffic),
        "traffic_note": None if traffic is None else "committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this "
                                                     "workload (profiles/), not measured in this run",
        "kernel": name,
        "avg_launch_us": round(1e3 * p["ms"] / p["launches"], 2),
        "alg_bytes_per_launch": round(alg),
        "kernel_ms_total": round(p["ms"], 3),
    }


def kernel_table(ctx, kernel_ids):
    out = {}
    for k in kernel_ids:
        p = ctx.profile(k)
        if p["launches"]:
            out[KERNEL_NAMES[k]] = {"launches": int(p["launches"]), "ms": round(p["ms"], 3),
                                    "GB/s": round(p["alg_bytes"] / (p["ms"] * 1e-3) / 1e9, 1)}
    return out


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload`, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summaries (scaled by the counters' reading for a plain
    copy of known size), or None when there is none.  A PageRank pull "launch" is
    one dense round's pull launches."""
    for f, key in ((ROOT / "profiles" / f"{kernel}_{workload}_traffic.txt", "traffic_bytes_per_launch:"),
                   (ROOT / "profiles" / f"{kernel}_traffic.txt", f"traffic_bytes_per_launch[{workload}]:")):
        if f.exists():
            for line in f.read_text().splitlines():
                if line.startswith(key):
                    return float(line.split(":", 1)[1])
    return None


def run_sssp(args, wl):
    import oracle
    import suminc_amd
    import suminc_amd._lib as L

    sketch = args.supernodes == "on"
    ctx = suminc_amd.Context(max_rounds=1_000_000)
    t0 = time.time()
    ctx.generate_rmat(wl["scale"], 16, 0.57, 0.19, 0.19, seed=wl["seed"], weights="u32", wmin=wl["wmin"],
                      wmax=wl["wmax"])
    n, m, _ = ctx.graph_info()
    log(f"graph: n={n} m={m} ({time.time() - t0:.1f}s)")
    extra = {}
    if sketch:
        t0 = time.time()
        ctx.cluster_leaves(wl["max_node"], 2)
        info = ctx.sum_build(2, wl["max_node"], 1.0)
        log(f"sketch: {info} ({time.time() - t0:.1f}s)")
        extra["sketch"] = {k: info[k] for k in ("clusters", "clustered_vertices", "entries", "sketch_edges",
                                                "internal_edges", "shortcut_slots", "max_cluster")}
        extra["sketch"]["node_type_count"] = info["node_type_count"]
    run = ctx.sum_sssp_run if sketch else ctx.sssp_run
    ctx.sssp_init(wl["source"])
    st = run()
    log(f"initial SSSP: rounds={st['rounds']:.0f} edges={st['edges_traversed']:.3e} {st['kernel_ms']:.1f} ms")
    n_ops = int(wl["frac"] * m)
    batches = [ctx.generate_batch(wl["seed"], b, n_ops // 2, n_ops - n_ops // 2, wl["wmin"], wl["wmax"])
               for b in range(args.warmup + args.steps)]
    dev = [ctx.upload_batch(b) for b in batches]  # inputs resident in HBM before the timed region
    for b in range(args.warmup):
        ctx.apply_batch_device(dev[b])
        run()
    ctx.sync()
    ctx.profile_enable(True)
    edges = 0.0
    run_ms = 0.0
    t0 = time.perf_counter()
    for b in range(args.warmup, args.warmup + args.steps):
        ctx.apply_batch_device(dev[b])
        st = run()
        edges += st["edges_traversed"]
        run_ms += st["kernel_ms"]
    ctx.sync()
    dt = time.perf_counter() - t0
    kids = [L.KERNEL_SSSP_SKETCH if sketch else L.KERNEL_SSSP_PUSH, L.KERNEL_BATCH_COPY]
    kernel, roof = roofline(ctx, kids, args.workload)
    extra["kernels"] = kernel_table(ctx, kids)
    kernel = name_pull(args, ctx.graph_info()[0], kernel, roof, extra["kernels"])
    ctx.profile_enable(False)
    for d in dev:
        d.free()
    ms_per_step = 1e3 * dt / args.steps
    gteps = edges / dt / 1e9
    log(f"incremental: {ms_per_step:.3f} ms/step (query {run_ms / args.steps:.3f} ms), "
        f"{edges / args.steps:.3e} edges/step, {gteps:.4f} GTEPS")

    # host-buffer variant of the same step (PCIe upload of the ops inside the step),
    # on further fresh batches of the same size
    n_host = min(args.steps, 5)
    host_batches = [ctx.generate_batch(wl["seed"], b, n_ops // 2, n_ops - n_ops // 2, wl["wmin"], wl["wmax"])
                    for b in range(args.warmup + args.steps, args.warmup + args.steps + n_host)]
    ctx.sync()
    t1 = time.perf_counter()
    for hb in host_batches:
        ctx.apply_batch(hb)
        run()
    ctx.sync()
    host_ms = 1e3 * (time.perf_counter() - t1) / n_host

    # the incremental state is what gets checked
    off, col, w = ctx.download_csr()
    dist, par = ctx.sssp_download()

    # full recompute on the same GPU, same final graph, plain path
    full = []
    for _ in range(3):
        ctx.sync()
        t1 = time.perf_counter()
        ctx.sssp_init(wl["source"])
        stf = ctx.sssp_run()
        ctx.sync()
        full.append(time.perf_counter() - t1)
    full_ms = 1e3 * float(np.median(full))
    log(f"full recompute: {full_ms:.3f} ms ({stf['edges_traversed']:.3e} edges)")

    extra.update({"full_recompute_ms": round(full_ms, 3), "speedup_vs_full_recompute": round(full_ms / ms_per_step, 2),
                  "query_ms_per_step": round(run_ms / args.steps, 4), "edges_per_step": edges / args.steps,
                  "host_ops_ms_per_step": round(host_ms, 4)})
    cpu = None
    if not args.no_verify or not args.no_cpu:
        g = oracle.Graph.from_csr(off, col, w)
        cores = min(host_cores(), 16)
        oracle.set_threads(cores)
        if not args.no_verify:
            t1 = time.time()
            od, op = g.dijkstra(wl["source"])
            extra["bit_exact_dist"] = bool(np.array_equal(dist, od))
            extra["bit_exact_parent"] = bool(np.array_equal(par, op))
            log(f"verify vs Dijkstra ({time.time() - t1:.1f}s): dist {extra['bit_exact_dist']} "
                f"parent {extra['bit_exact_parent']}")
        if not args.no_cpu:
            # one more seeded batch of the same shape, applied on the host (untimed),
            # then the reference's incremental step restated on the CPU, repeated
            ops = g.gen_batch(wl["seed"], 10_000, n_ops // 2, n_ops - n_ops // 2, wl["wmin"], wl["wmax"])
            t1 = time.time()
            g2, _ = g.apply_batch(ops)
            log(f"cpu baseline: host graph update {time.time() - t1:.1f}s (untimed)")
            reps, tc, ec = 0, 0.0, 0.0
            while tc < args.cpu_seconds and reps < 1000:
                t1 = time.perf_counter()
                _, _, sto = g2.sssp_inc(g, wl["source"], dist, par)
                tc += time.perf_counter() - t1
                ec += sto["edges"]
                reps += 1
            cpu = {"value": round(ec / tc / 1e9, 4), "unit": "GTEPS", "cores": cores, "kind": "port",
                   "ms_per_step": round(1e3 * tc / reps, 3),
                   "sample": f"oracle restatement of the reference's incremental SSSP step (deltaCompute reset + "
                             f"resend + Ingress rounds) on the s{wl['scale']} graph after one more "
                             f"{100 * wl['frac']:g}% batch, repeated {reps}x: {ec:.3e} edges in {tc:.1f}s "
                             f"(host graph update excluded)"}
            log(f"cpu baseline: {cpu}")
    ctx.close()
    return gteps, ms_per_step, "u32", roof, cpu, extra


def run_pr(args, wl):
    import oracle
    import suminc_amd
    import suminc_amd._lib as L

    ctx = suminc_amd.Context(termcheck_threshold=wl["T"], pr_precision=suminc_amd.PR_FP32_KAHAN,
                             max_rounds=1_000_000, **pull_kw(args))
    t0 = time.time()
    ctx.generate_rmat(wl["scale"], 16, 0.57, 0.19, 0.19, seed=wl["seed"])
    n, m, _ = ctx.graph_info()
    log(f"graph: n={n} m={m} ({time.time() - t0:.1f}s)")
    sketch = args.supernodes == "on" and wl.get("max_node")
    run = ctx.sum_pr_run if sketch else ctx.pr_run
    extra_sk = {}
    if sketch:
        t0 = time.time()
        ctx.cluster_leaves(wl["max_node"], 2)
        info = ctx.sum_build(2, wl["max_node"], 1.0)
        log(f"sketch: {info} ({time.time() - t0:.1f}s)")
        extra_sk = {k: info[k] for k in ("clusters", "clustered_vertices", "entries", "sketch_edges",
                                         "internal_edges", "shortcut_slots", "max_cluster")}
    ctx.pr_init()
    st = run()
    log(f"initial PR: rounds={st['rounds']:.0f} edges={st['edges_traversed']:.3e} {st['kernel_ms']:.1f} ms")
    n_ops = int(wl["frac"] * m)
    batches = [ctx.generate_batch(wl["seed"], b, n_ops // 2, n_ops - n_ops // 2)
               for b in range(args.warmup + args.steps)]
    for b in range(args.warmup):
        ctx.apply_batch(batches[b])
        run()
    ctx.sync()
    ctx.profile_enable(True)
    edges = 0.0
    t0 = time.perf_counter()
    for b in range(args.warmup, args.warmup + args.steps):
        sa = ctx.apply_batch(batches[b])
        st = run()
        edges += st["edges_traversed"] + sa["reseed_edges"]
    ctx.sync()
    dt = time.perf_counter() - t0
    kids = [L.KERNEL_PR_PUSH, L.KERNEL_PR_PULL, L.KERNEL_PR_GATHER, L.KERNEL_BATCH_COPY]
    kernel, roof = roofline(ctx, kids, args.workload)
    ktab = kernel_table(ctx, kids)
    kernel = name_pull(args, ctx.graph_info()[0], kernel, roof, ktab)
    if roof is not None and kernel != "k_pr_pull":
        # the committed counter summaries are the row-class pull's: the binned and
        # tiled pulls are measured now, in their own counter runs after this one
        roof["traffic"] = None
        roof["traffic_note"] = None if not args.no_traffic else f"no counter run for {kernel} (--no-traffic)"
    ctx.profile_enable(False)
    ms_per_step = 1e3 * dt / args.steps
    gteps = edges / dt / 1e9
    log(f"incremental: {ms_per_step:.3f} ms/step, {edges / args.steps:.3e} edges/step, {gteps:.3f} GTEPS")
    rank, _ = ctx.pr_download()
    full = []
    for _ in range(2):
        ctx.sync()
        t1 = time.perf_counter()
        ctx.pr_init()
        stf = ctx.pr_run()
        ctx.sync()
        full.append(time.perf_counter() - t1)
    full_ms = 1e3 * float(np.median(full))
    extra = {"full_recompute_ms": round(full_ms, 3), "speedup_vs_full_recompute": round(full_ms / ms_per_step, 2),
             "edges_per_step": edges / args.steps, "termcheck_threshold": wl["T"],
             "kernels": ktab}
    if sketch:
        extra["sketch"] = extra_sk
    cpu = None
    if not args.no_verify or not args.no_cpu:
        off, col, _ = ctx.download_csr()
        g = oracle.Graph.from_csr(off, col)
        cores = min(host_cores(), 16)
        oracle.set_threads(cores)
        if not args.no_verify:
            rs, _, _ = g.pr_solve(DAMPING, tol=1e-13)
            extra["l1_vs_fp64"] = float(np.abs(rank - rs).sum())
            extra["linf_vs_fp64"] = float(np.abs(rank - rs).max())
            log(f"L1 vs FP64 = {extra['l1_vs_fp64']:.3e}")
        if not args.no_cpu:
            t1 = time.perf_counter()
            _, _, sto = g.pr_ingress(DAMPING, wl["T"], max_rounds=30)
            tc = time.perf_counter() - t1
            cpu = {"value": round(sto["edges"] / tc / 1e9, 4), "unit": "GTEPS", "cores": cores, "kind": "port",
                   "sample": f"oracle delta-PageRank Ingress (reference CPU loop restated, FP64) from scratch on "
                             f"the final s{wl['scale']} graph, first {sto['rounds']:.0f} rounds: "
                             f"{sto['edges']:.3e} edges in {tc:.2f}s"}
            log(f"cpu baseline: {cpu}")
    ctx.close()
    if roof is not None and kernel != "k_pr_pull" and not args.no_traffic:
        live_traffic(args, kernel, roof)
    return gteps, ms_per_step, "f32", roof, cpu, extra


def _pmc_traffic(args, kernel_subs):
    """HBM traffic per round of the kernels whose rocprof names contain the strings
    of kernel_subs (one launch of each per round), measured now: two rocprofv3 --pmc
    runs of a short child of this script (FETCH_SIZE, then WRITE_SIZE, KiB per
    dispatch), each scaled by what the same counter reads for a plain
    device-to-device copy of known size (FETCH x2, WRITE x1 on MI355X,
    profiles/k_copy_old_sssp22_traffic.txt), the per-dispatch means of the kernels
    added.  Returns (bytes, note) or (None, reason)."""
    import glob
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3")
                                         else None)
    if prof is None:
        return None, "rocprofv3 not found"
    scale = {"FETCH_SIZE": 2.0, "WRITE_SIZE": 1.0}
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix=f"pmc_{counter.lower()}_", dir=os.environ.get("TMPDIR", "/tmp"))
        cmd = ["timeout", "-s", "KILL", str(args.traffic_timeout), prof, "--pmc", counter, "--output-format", "csv",
               "-d", d, "-o", "run", "--", sys.executable, str(Path(__file__).resolve()), "--traffic-child",
               "--workload", args.workload, "--pull-band-kib", str(args.pull_band_kib),
               "--pull-tile-kib", str(args.pull_tile_kib), "--pull-bin-kib", str(args.pull_bin_kib),
               "--pull-bin-direct", str(args.pull_bin_direct)]
        log(f"traffic: {' '.join(cmd[4:])}")
        r = subprocess.run(cmd, stdout=sys.stderr, stderr=sys.stderr, cwd=str(ROOT))  # progress passes through
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            shutil.rmtree(d, ignore_errors=True)
            return None, f"rocprofv3 --pmc {counter} wrote no counter file (rc {r.returncode})"
        vals = {k: [] for k in kernel_subs}
        import csv

        for f in files:
            with open(f, newline="") as fh:
                for row in csv.DictReader(fh):
                    name = row.get("Kernel_Name") or row.get("Kernel-Name") or ""
                    for k in kernel_subs:
                        if k in name and row.get("Counter_Name", counter) == counter:
                            vals[k].append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        mean_kib, count = 0.0, 0
        for k, v in vals.items():
            if not v:
                return None, f"no {k} dispatch in the {counter} run"
            top = max(v)
            work = [x for x in v if x >= 0.1 * top]  # dispatches that did work (idle ones return at entry)
            mean_kib += float(np.mean(work))
            count = len(work)
        out[counter] = (mean_kib, count)
    b = sum(out[k][0] * 1024.0 * scale[k] for k in out)
    note = (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over {out['FETCH_SIZE'][1]} dispatches of "
            f"{' + '.join(kernel_subs)} in a {args.workload} child run: FETCH {out['FETCH_SIZE'][0]:.0f} KiB x2, "
            f"WRITE {out['WRITE_SIZE'][0]:.0f} KiB x1 per round")
    return b, note


def live_traffic(args, kernel, roof):
    """roofline.traffic of the FP32 PageRank pull kernels, measured now by two counter
    runs of the traffic child (_pmc_traffic)."""
    sub = {"k_pr_pull": ["k_pr_pull<float>"], "k_pr_pull_tiled": ["k_pr_pull_tiled<float>"],
           PULL_KERNELS["binned"]: ["k_pr_bin_gather<float, false>", "k_pr_bin_sum<float, false>"]}.get(kernel)
    if sub is None:
        roof["traffic_note"] = f"no live counter run for {kernel}"
        return
    tb, note = _pmc_traffic(args, sub)
    roof["traffic"] = None if tb is None else round(tb)
    roof["traffic_note"] = note
    if tb is not None:
        roof["traffic_over_algorithmic"] = round(tb / roof["alg_bytes_per_launch"], 3)


def traffic_child(args, wl):
    """The short run the traffic measurement profiles: the same graph, sketch and
    precision as the workload, one SumInc PageRank query from the initial state cut
    at a few rounds per phase (every round from the initial state is a dense pull)."""
    import suminc_amd

    ctx = suminc_amd.Context(termcheck_threshold=wl["T"], pr_precision=suminc_amd.PR_FP32_KAHAN, max_rounds=4,
                             **pull_kw(args))
    ctx.generate_rmat(wl["scale"], 16, 0.57, 0.19, 0.19, seed=wl["seed"])
    ctx.cluster_leaves(wl["max_node"], 2)
    ctx.sum_build(2, wl["max_node"], 1.0)
    ctx.pr_init()
    try:
        ctx.sum_pr_run()
    except suminc_amd.SumincError as e:
        if e.code != suminc_amd.E_NOT_CONVERGED:
            raise
    ctx.pr_init()
    ctx.pr_run(allow_not_converged=True)
    ctx.sync()
    ctx.close()


def run_web(args, wl):
    """The web-like incremental run (s26, successive 1% batches).  A step applies one
    batch from HBM (CSR/CSC merge, PageRank reseed, SSSP repair, sketch bookkeeping)
    and brings both states back to their fixed points: the SumInc PageRank query
    over the supernode sketch (precompute, sketch rounds, correction, residual
    rounds) and the SumInc SSSP query.  Beside it: the same batches with PageRank on
    the plain CSR, full recomputes on the same GPU, SSSP against a plain recompute
    bit for bit, PageRank against an FP64 run (L1, Linf), the weighted-PageRank
    forward + adjoint, and the reference's incremental step restated on the CPU."""
    import oracle
    import suminc_amd
    import suminc_amd._lib as L

    T = wl["T"]
    ctx = suminc_amd.Context(termcheck_threshold=T, pr_precision=suminc_amd.PR_FP32_KAHAN,
                             max_rounds=1_000_000, **pull_kw(args))
    t0 = time.time()
    ctx.generate_rmat(wl["scale"], 16, 0.57, 0.19, 0.19, seed=wl["seed"], weights="u32", wmin=wl["wmin"],
                      wmax=wl["wmax"])
    n, m, _ = ctx.graph_info()
    log(f"graph: n={n} m={m} ({time.time() - t0:.1f}s)")
    t0 = time.time()
    ctx.cluster_leaves(wl["max_node"], 2)
    info = ctx.sum_build(2, wl["max_node"], 1.0)
    log(f"sketch: {info} ({time.time() - t0:.1f}s)")
    sk = {k: info[k] for k in ("clusters", "clustered_vertices", "entries", "sketch_edges", "internal_edges",
                               "shortcut_slots", "max_cluster")}
    sk["node_type_count"] = info["node_type_count"]
    # edges the sketch rounds carry: Singles' and exits' external out-edges plus one
    # shortcut slot per (entry, exit) pair, against the plain graph's m
    sk["compression_ratio"] = round((info["sketch_edges"] + info["shortcut_slots"]) / m, 4)
    extra = {"sketch": sk}
    ctx.pr_init()
    t0 = time.time()
    st = ctx.sum_pr_run()
    log(f"initial PR (SumInc query): rounds={st['rounds']:.0f} ({time.time() - t0:.1f}s)")
    ctx.sssp_init(wl["source"])
    st = ctx.sum_sssp_run()
    log(f"initial SSSP: rounds={st['rounds']:.0f} edges={st['edges_traversed']:.3e}")
    n_ops = int(wl["frac"] * m)
    n_plain = 3
    nb = args.warmup + args.steps + n_plain
    dev = []
    for b in range(nb):  # inputs resident in HBM before the timed region
        ops = ctx.generate_batch(wl["seed"], b, n_ops // 2, n_ops - n_ops // 2, wl["wmin"], wl["wmax"])
        dev.append(ctx.upload_batch(ops))
    del ops
    for b in range(args.warmup):
        ctx.apply_batch_device(dev[b])
        ctx.sum_pr_run()
        ctx.sum_sssp_run()
    ctx.sync()
    ctx.profile_enable(True)
    per = []
    edges = 0.0
    t_all = time.perf_counter()
    pull_prev = ctx.profile(L.KERNEL_PR_PULL)
    for b in range(args.warmup, args.warmup + args.steps):
        t0 = time.perf_counter()
        sa = ctx.apply_batch_device(dev[b])
        ctx.sync()
        t1 = time.perf_counter()
        sp = ctx.sum_pr_run()
        ctx.sync()
        t2 = time.perf_counter()
        pull_cur = ctx.profile(L.KERNEL_PR_PULL)
        n_pull = pull_cur["launches"] - pull_prev["launches"]
        pull_ms = (pull_cur["ms"] - pull_prev["ms"]) / n_pull if n_pull else 0.0
        pull_prev = pull_cur
        ss = ctx.sum_sssp_run()
        ctx.sync()
        t3 = time.perf_counter()
        e = sp["edges_traversed"] + sa["reseed_edges"] + ss["edges_traversed"]
        edges += e
        per.append({"batch": b, "ms": round(1e3 * (t3 - t0), 3), "apply_ms": round(1e3 * (t1 - t0), 3),
                    "pr_ms": round(1e3 * (t2 - t1), 3), "sssp_ms": round(1e3 * (t3 - t2), 3),
                    "pr_sketch_rounds": sp["sketch_rounds"], "pr_residual_rounds": sp["residual_rounds"],
                    "sssp_rounds": ss["rounds"], "mass_defect": sp["mass_defect"], "edges": e,
                    "pull_ms_per_round": round(pull_ms, 3),
                    "GTEPS": round(e / (t3 - t0) / 1e9, 3)})
        log(f"batch {b}: {per[-1]}")
    ctx.sync()
    dt = time.perf_counter() - t_all
    kids = [L.KERNEL_PR_PUSH, L.KERNEL_PR_PULL, L.KERNEL_PR_GATHER, L.KERNEL_SSSP_SKETCH, L.KERNEL_BATCH_COPY]
    kernel, roof = roofline(ctx, kids, args.workload)
    extra["kernels"] = kernel_table(ctx, kids)
    kernel = name_pull(args, ctx.graph_info()[0], kernel, roof, extra["kernels"])
    ctx.profile_enable(False)
    ms_per_step = 1e3 * dt / args.steps
    gteps = edges / dt / 1e9
    log(f"incremental: {ms_per_step:.3f} ms/step, {edges / args.steps:.3e} edges/step, {gteps:.3f} GTEPS")
    extra["batches"] = per

    # checks on the graph of the last step: SSSP against a plain full recompute bit
    # for bit, PageRank against an FP64 run (L1, Linf), and the full recomputes' time
    rank, delta = ctx.pr_download()
    dist, par = ctx.sssp_download()
    t0 = time.perf_counter()
    ctx.sssp_init(wl["source"])
    ctx.sssp_run()
    ctx.sync()
    full_ss = 1e3 * (time.perf_counter() - t0)
    d2, p2 = ctx.sssp_download()
    extra["sssp_equals_full_recompute"] = bool(np.array_equal(dist, d2) and np.array_equal(par, p2))
    del d2, p2
    t0 = time.perf_counter()
    ctx.pr_init()
    stf = ctx.pr_run()
    ctx.sync()
    full_pr = 1e3 * (time.perf_counter() - t0)
    mean = lambda xs, k: float(np.mean([x[k] for x in xs]))  # noqa: E731
    extra["mass_defect_max"] = float(max(abs(x["mass_defect"]) for x in per))
    if not args.no_verify:
        ctx.set_params(pr_precision=suminc_amd.PR_FP64, termcheck_threshold=1e-12)
        ctx.pr_init()
        st = ctx.pr_run()
        r64, _ = ctx.pr_download()
        extra["l1_vs_fp64"] = float(np.abs(rank - r64).sum())
        extra["linf_vs_fp64"] = float(np.abs(rank - r64).max())
        extra["fp64_reference"] = f"plain FP64 delta-PageRank on the same GPU and graph to sum|delta| < 1e-12 " \
                                  f"({st['rounds']:.0f} rounds, mass defect {st['mass_defect']:.2e})"
        log(f"L1 vs FP64 = {extra['l1_vs_fp64']:.3e}, Linf {extra['linf_vs_fp64']:.3e}")
        del r64
        ctx.set_params(pr_precision=suminc_amd.PR_FP32_KAHAN, termcheck_threshold=T)
    ctx.pr_upload(rank, delta)  # back to the incremental state

    # the same batch size with PageRank on the plain CSR (SSSP as in the steps)
    plain = []
    for b in range(args.warmup + args.steps, nb):
        t0 = time.perf_counter()
        ctx.apply_batch_device(dev[b])
        ctx.sync()
        t1 = time.perf_counter()
        sp = ctx.pr_run()
        ctx.sync()
        t2 = time.perf_counter()
        ctx.sum_sssp_run()
        plain.append({"apply_ms": 1e3 * (t1 - t0), "pr_ms": 1e3 * (t2 - t1), "rounds": sp["rounds"]})
    for d in dev:
        d.free()
    extra["plain_csr_pagerank"] = {"batches": n_plain, "pr_ms": round(mean(plain, "pr_ms"), 3),
                                   "apply_ms": round(mean(plain, "apply_ms"), 3),
                                   "rounds": round(mean(plain, "rounds"), 1)}
    log(f"plain CSR PageRank: {extra['plain_csr_pagerank']}")
    inc_pr = mean(per, "pr_ms") + mean(per, "apply_ms")
    inc_plain = mean(plain, "pr_ms") + mean(plain, "apply_ms")
    extra.update({"full_recompute_ms": {"pagerank": round(full_pr, 3), "pagerank_rounds": stf["rounds"],
                                        "sssp": round(full_ss, 3)},
                  "incremental_ms": {"pagerank_sketch": round(inc_pr, 3), "pagerank_plain": round(inc_plain, 3),
                                     "sssp": round(mean(per, "sssp_ms"), 3)},
                  "speedup_vs_full_recompute": {
                      "pagerank_sketch": round(full_pr / inc_pr, 2),
                      "pagerank_plain": round(full_pr / inc_plain, 2),
                      "sssp": round(full_ss / mean(per, "sssp_ms"), 2),
                      "step": round((full_pr + full_ss) / ms_per_step, 2)},
                  "edges_per_step": edges / args.steps, "termcheck_threshold": T})
    log(f"full recompute: pagerank {full_pr:.1f} ms, sssp {full_ss:.1f} ms; sssp bit-exact vs full: "
        f"{extra['sssp_equals_full_recompute']}")
    rank, delta = ctx.pr_download()
    dist, par = ctx.sssp_download()
    cpu = None
    if not args.no_cpu:
        cpu = cpu_incremental_step(args, wl, ctx, rank, delta, dist, par, mean(plain, "rounds"))
    ctx.close()
    del rank, delta, dist, par
    if not args.no_adjoint:
        # weighted personalized PageRank + adjoint on the same R-MAT graph with FP32
        # weights (a context of its own: the SSSP graph's weights are integers),
        # Dirichlet(1) personalization over random seeds, c ~ N(0,1)
        with suminc_amd.Context(termcheck_threshold=wl["T"], pr_precision=suminc_amd.PR_FP32_KAHAN,
                                max_rounds=1_000_000, **pull_kw(args)) as cw:
            cw.generate_rmat(wl["scale"], 16, 0.57, 0.19, 0.19, seed=wl["seed"], weights="f32")
            rng = np.random.default_rng(wl["seed"])
            pvec = np.zeros(n, np.float32)
            seeds = rng.choice(n, wl["adjoint_seeds"], replace=False)
            pvec[seeds] = rng.dirichlet(np.ones(wl["adjoint_seeds"])).astype(np.float32)
            cvec = rng.standard_normal(n, dtype=np.float32)
            _, sf = cw.wpr_forward(pvec)
            _, gw, sb = cw.wpr_backward(cvec)
            mw = cw.graph_info()[1]
            extra["adjoint"] = {"edges": mw, "forward_ms": round(sf["kernel_ms"], 3),
                                "forward_rounds": sf["rounds"], "backward_ms": round(sb["kernel_ms"], 3),
                                "backward_rounds": sb["rounds"], "grad_w_finite": bool(np.isfinite(gw).all())}
            log(f"adjoint: {extra['adjoint']}")
            del gw
    if roof is not None and not args.no_traffic:
        live_traffic(args, kernel, roof)
    return gteps, ms_per_step, "f32+u32", roof, cpu, extra


def cpu_incremental_step(args, wl, ctx, rank, delta, dist, par, plain_rounds):
    """The reference's incremental step restated on the host (oracle, FP64): the
    graph update (IncFragmentBuilder's rebuild), AmendValue(-1) over every vertex on
    the old graph and AmendValue(+1) on the new one (sum_sync_iter_worker.h:139-345,
    the full sweeps the reference runs), Ingress rounds (:857-882), and the
    incremental SSSP step (deltaCompute reset + resend + rounds,
    sum_sync_traversal_worker.h:113-352).  All cores, then one thread, on a bounded
    sample: the PageRank rounds stop after about cpu_seconds."""
    import oracle

    log("cpu baseline: downloading the final graph")
    off, col, w = ctx.download_csr()
    n = len(off) - 1
    m = len(col)
    g = oracle.Graph.from_csr(off, col, w)
    del off, col, w
    # one more batch of the same size; the updated graph comes from the device's
    # merge (the reference rebuilds its fragment on the host, which is not timed here)
    n_ops = int(wl["frac"] * m)
    ops = ctx.generate_batch(wl["seed"], 10_000, n_ops // 2, n_ops - n_ops // 2, wl["wmin"], wl["wmax"])
    ctx.apply_batch(ops)
    log("cpu baseline: downloading the updated graph")
    off, col, w = ctx.download_csr()
    g2 = oracle.Graph.from_csr(off, col, w)
    del off, col, w
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or host_cores()
    nproc = os.cpu_count()
    res = {}
    for label, th in (("all", threads), ("single", 1)):
        oracle.set_threads(th)
        r = rank.copy()
        dl = delta.copy()
        sample = 1.0 if label == "all" else 0.125  # one thread amends a seeded eighth of the vertices
        touched = None
        if sample < 1.0:
            touched = (np.random.default_rng(wl["seed"]).random(n) < sample).astype(np.uint8)
        t0 = time.perf_counter()
        g.pr_amend(DAMPING, -1, r, dl, touched)
        g2.pr_amend(DAMPING, +1, r, dl, touched)
        t_amend = time.perf_counter() - t0
        amend_edges = sample * (g.m + g2.m)
        log(f"cpu baseline ({label}, {th} threads): amend {t_amend:.1f}s")
        t0 = time.perf_counter()
        rounds, redges = 0, 0.0
        budget = args.cpu_seconds if label == "all" else args.cpu_seconds / 2
        while time.perf_counter() - t0 < budget and rounds < plain_rounds:
            r, dl, sto = g2.pr_ingress(DAMPING, wl["T"], max_rounds=1, r=r, delta=dl)
            rounds += 1
            redges += sto["edges"]
            log(f"cpu baseline ({label}): round {rounds} {time.perf_counter() - t0:.1f}s")
        t_rounds = time.perf_counter() - t0
        t_ss, ss_edges = 0.0, 0.0
        if label == "all":
            t0 = time.perf_counter()
            _, _, sts = g2.sssp_inc(g, wl["source"], dist, par)
            t_ss = time.perf_counter() - t0
            ss_edges = sts["edges"]
            log(f"cpu baseline ({label}): sssp step {t_ss:.1f}s")
        t_tot = t_amend + t_rounds + t_ss
        e_tot = amend_edges + redges + ss_edges
        per_round = t_rounds / max(rounds, 1)
        res[label] = {"threads": th, "GTEPS": round(e_tot / t_tot / 1e9, 4),
                      "amend_s": round(t_amend, 3), "amend_sample": sample, "pr_rounds_run": rounds,
                      "pr_round_s": round(per_round, 3), "sssp_inc_s": round(t_ss, 3),
                      "edges": e_tot}
        if label == "all":
            res[label]["ms_per_step_estimate"] = round(1e3 * (t_amend + per_round * plain_rounds + t_ss), 1)
        log(f"cpu baseline ({label}): {res[label]}")
    del g, g2
    oracle.set_threads(threads)
    a = res["all"]
    return {"value": a["GTEPS"], "unit": "GTEPS", "cores": threads, "kind": "port",
            "nproc": nproc, "affinity_cpus": host_cores(),
            "single_thread": res["single"], "all_cores": a,
            "ms_per_step_estimate": a["ms_per_step_estimate"],
            "sample": f"oracle restatement (FP64, OpenMP) of the reference's incremental step on the final "
                      f"s{wl['scale']} graph after one more {100 * wl['frac']:g}% batch: AmendValue(-1/+1) "
                      f"sweeps over every vertex, {a['pr_rounds_run']} Ingress rounds (the GPU's plain path "
                      f"needs {plain_rounds:.0f}; ms_per_step_estimate scales the measured round time to that), "
                      f"the incremental SSSP step; GTEPS = edges amended/pushed/relaxed per second; one thread: "
                      f"amend over a seeded eighth of the vertices plus {res['single']['pr_rounds_run']} rounds; "
                      f"the host fragment rebuild is not in the sample (the updated graph comes from the device)"}


def run_prfull(args, wl):
    """The reference's CPU-runnable case on the device: plain FP64 delta-PageRank
    from scratch, each step a full run from the initial state; the oracle's
    restatement of the reference's CPU loop is timed on the same graph."""
    import oracle
    import suminc_amd
    import suminc_amd._lib as L

    g = oracle.Graph.rmat(wl["scale"], 16, seed=wl["seed"])
    off, col, _ = g.csr()
    T = wl["eps"] * g.n
    ctx = suminc_amd.Context(termcheck_threshold=T, pr_precision=suminc_amd.PR_FP64, max_rounds=1_000_000)
    ctx.upload_csr(off, col)
    for _ in range(args.warmup):
        ctx.pr_init()
        ctx.pr_run()
    ctx.sync()
    ctx.profile_enable(True)
    edges = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.pr_init()
        st = ctx.pr_run()
        edges += st["edges_traversed"]
    ctx.sync()
    dt = time.perf_counter() - t0
    kids = [L.KERNEL_PR_PUSH, L.KERNEL_PR_PULL, L.KERNEL_PR_GATHER]
    kernel, roof = roofline(ctx, kids, args.workload)
    extra = {"kernels": kernel_table(ctx, kids), "rounds": st["rounds"], "termcheck_threshold": T}
    kernel = name_pull(args, ctx.graph_info()[0], kernel, roof, extra["kernels"])
    ctx.profile_enable(False)
    r, _ = ctx.pr_download()
    ctx.close()
    ms_per_step = 1e3 * dt / args.steps
    log(f"full run: {ms_per_step:.3f} ms ({st['rounds']:.0f} rounds)")
    if not args.no_verify:
        rs, _, _ = g.pr_solve(DAMPING, tol=1e-15)
        extra["l1_vs_fp64_fixed_point"] = float(np.abs(r - rs).sum())
        extra["l1_bound"] = st["residual"] / (1 - DAMPING)
    cpu = None
    if not args.no_cpu:
        cores = min(host_cores(), 16)
        oracle.set_threads(cores)
        t1 = time.perf_counter()
        ro, _, sto = g.pr_ingress(DAMPING, T, max_rounds=100000)
        tc = time.perf_counter() - t1
        cpu = {"value": round(sto["edges"] / tc / 1e9, 4), "unit": "GTEPS", "cores": cores, "kind": "port",
               "ms_per_run": round(1e3 * tc, 3),
               "sample": f"oracle delta-PageRank Ingress (reference CPU loop restated, FP64) to the same stop rule: "
                         f"{sto['rounds']:.0f} rounds, {sto['edges']:.3e} edges in {tc:.2f}s"}
        extra["cpu_l1_vs_gpu"] = float(np.abs(ro - r).sum())
        log(f"cpu baseline: {cpu}")
    return edges / dt / 1e9, ms_per_step, "f64", roof, cpu, extra


def run_wpr(args, wl):
    """Forward + adjoint of the weighted personalized PageRank, timed per step;
    checked against an FP64 run of both on the same GPU and against central
    differences of the FP64 forward on random edges."""
    import oracle
    import suminc_amd

    rng = np.random.default_rng(wl["seed"])
    ctx = suminc_amd.Context(termcheck_threshold=wl["T"], pr_precision=suminc_amd.PR_FP32_KAHAN,
                             max_rounds=1_000_000)
    t0 = time.time()
    ctx.generate_rmat(wl["scale"], 16, 0.57, 0.19, 0.19, seed=wl["seed"], weights="f32")
    n, m, _ = ctx.graph_info()
    log(f"graph: n={n} m={m} ({time.time() - t0:.1f}s)")
    p = np.zeros(n, np.float32)
    seeds = rng.choice(n, wl["seeds"], replace=False)
    p[seeds] = rng.dirichlet(np.ones(wl["seeds"])).astype(np.float32)
    cvec = rng.standard_normal(n, dtype=np.float32)
    for _ in range(args.warmup):
        ctx.wpr_forward(p)
        ctx.wpr_backward(cvec)
    ctx.sync()
    edges, kms, rounds = 0.0, 0.0, 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        _, sf = ctx.wpr_forward(p)
        gp, gw, sb = ctx.wpr_backward(cvec)
        edges += sf["edges_traversed"] + sb["edges_traversed"]
        kms += sf["kernel_ms"] + sb["kernel_ms"]
        rounds += sf["rounds"] + sb["rounds"]
    ctx.sync()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    gteps = edges / dt / 1e9
    log(f"forward+adjoint: {ms_per_step:.3f} ms/step ({rounds / args.steps:.0f} rounds), {gteps:.3f} GTEPS")
    # roofline of the series rounds (row walk, coefficient and gathered value per
    # edge; previous value, new value and the compensated sum per vertex)
    es = 4.0
    alg = rounds * (m * (4.0 + 2 * es) + n * (4 * es + 4.0))
    achieved = alg / (kms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": ("k_pr_bin_gather+k_pr_bin_sum (with coefficients)" if pull_mode(args, n) == "binned"
                       else "k_spmv") + " + k_wpr_step (series rounds)",
            "avg_launch_us": round(1e3 * kms / rounds, 2),
            "alg_bytes_per_launch": round(alg / rounds), "kernel_ms_total": round(kms, 3)}
    r32, _ = ctx.wpr_download()
    off, col, w = ctx.download_csr()
    ctx.close()
    extra = {"rounds_per_step": rounds / args.steps}
    if not args.no_verify:
        h = wl["h"]
        with suminc_amd.Context(termcheck_threshold=1e-15, pr_precision=suminc_amd.PR_FP64,
                                max_rounds=1_000_000) as c64:
            c64.upload_csr(off, col, w)
            c64.wpr_forward(p)
            r64, _ = c64.wpr_download()
            gp64, gw64, _ = c64.wpr_backward(cvec)
            extra["forward_l1_vs_fp64"] = float(np.abs(r32 - r64).sum())
            extra["grad_w_max_rel_vs_fp64"] = float(np.abs(gw - gw64).max() / np.abs(gw64).max())
            extra["grad_p_max_rel_vs_fp64"] = float(np.abs(gp - gp64).max() / np.abs(gp64).max())
            edges_fd = rng.choice(m, wl["fd_edges"], replace=False)
            fd = []
            t1 = time.time()
            for e in edges_fd:
                w0 = w[e]
                w[e] = w0 + np.float32(h)
                hp = np.float64(w[e])
                c64.wpr_set_weights(w)
                c64.wpr_forward(p)
                l1 = cvec.astype(np.float64) @ c64.wpr_download()[0]
                w[e] = w0 - np.float32(h)
                hm = np.float64(w[e])
                c64.wpr_set_weights(w)
                c64.wpr_forward(p)
                l2 = cvec.astype(np.float64) @ c64.wpr_download()[0]
                w[e] = w0
                fd.append((l1 - l2) / (hp - hm))
            fd = np.array(fd)
            ga = gw[edges_fd].astype(np.float64)
            rel = np.abs(fd - ga) / np.maximum(np.abs(fd), 1e-2 * np.abs(fd).max())
            extra["fd_check"] = {"edges": int(wl["fd_edges"]), "h": h, "max_rel_err": float(rel.max()),
                                 "within_1e-4": bool(rel.max() <= 1e-4),
                                 "reference": "central differences of the FP64 forward on the same GPU"}
            log(f"verify: {extra} ({time.time() - t1:.1f}s)")
    cpu = None
    if not args.no_cpu:
        g = oracle.Graph.from_csr(off, col, w)
        cores = min(host_cores(), 16)
        oracle.set_threads(cores)
        t1 = time.perf_counter()
        _, ro = g.wpr_forward(w, p.astype(np.float64), DAMPING, tol=1e-30, max_rounds=10)
        tc = time.perf_counter() - t1
        cpu = {"value": round(ro * m / tc / 1e9, 4), "unit": "GTEPS", "cores": cores, "kind": "port",
               "sample": f"oracle weighted-PageRank series (FP64, OpenMP) on the same s{wl['scale']} graph, "
                         f"first {ro} rounds: {ro * m:.3e} edges in {tc:.2f}s"}
        log(f"cpu baseline: {cpu}")
    return gteps, ms_per_step, "f32", roof, cpu, extra


def calibrate_copy(nbytes):
    """One device-to-device copy of nbytes (the runtime's copy kernel), so that the
    FETCH_SIZE / WRITE_SIZE counters can be read against a known byte count."""
    import ctypes as C

    import suminc_amd

    ctx = suminc_amd.Context()
    lib = ctx._lib
    a, b = C.c_void_p(), C.c_void_p()
    ctx._check(lib.suminc_device_alloc(ctx._h, nbytes, C.byref(a)))
    ctx._check(lib.suminc_device_alloc(ctx._h, nbytes, C.byref(b)))
    for _ in range(3):
        ctx._check(lib.suminc_copy_to_device(ctx._h, b, a, nbytes))
    ctx._check(lib.suminc_device_free(ctx._h, a))
    ctx._check(lib.suminc_device_free(ctx._h, b))
    ctx.close()
    log(f"copied {nbytes} bytes device to device, 3 times")


def supernodes_note(args, wl):
    if wl["kind"] in ("wpr", "prfull"):
        return "none (no batches, no sketch)"
    if wl["kind"] == "web":
        return ("PageRank and SSSP over supernodes (leaf clusters <= %d members); PageRank on the plain CSR "
                "reported beside it" % wl["max_node"])
    if args.supernodes == "on" and wl.get("max_node"):
        return "on (leaf clusters <= %d members, compress threshold 1)" % wl["max_node"]
    return "off (plain CSR path)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="web26", choices=sorted(WORKLOADS))
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--supernodes", default="on", choices=["on", "off"],
                    help="SSSP workload: query over the supernode sketch (on) or the plain CSR (off)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="length of the CPU baseline sample")
    ap.add_argument("--pull-band-kib", type=int, default=0,
                    help="PageRank workloads: source band of the banded pull in KiB (0: off, the library default)")
    ap.add_argument("--pull-tile-kib", type=int, default=0,
                    help="tiled pull source window in KiB (0: the library default, negative: the row-class pull)")
    ap.add_argument("--pull-bin-kib", type=int, default=0,
                    help="binned pull source window in KiB (0: the library default, negative: off)")
    ap.add_argument("--pull-bin-direct", type=int, default=-1,
                    help="binned pull: windows the summing pass gathers directly (negative: the library default)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES device-to-device (counter calibration for rocprofv3 --pmc) and exit")
    ap.add_argument("--no-adjoint", action="store_true", help="web26: skip the weighted-PageRank forward + adjoint")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure the dominant kernel's HBM traffic")
    ap.add_argument("--traffic-timeout", type=int, default=240, help="seconds per rocprofv3 counter run")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    KERNEL_NAMES.update(kernel_names())
    if args.traffic_child:
        traffic_child(args, WORKLOADS[args.workload])
        return
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU edge-cut path is designed, not built "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    wl = WORKLOADS[args.workload]
    if wl["kind"] == "sssp":
        value, ms, dtype, roof, cpu, extra = run_sssp(args, wl)
    elif wl["kind"] == "web":
        value, ms, dtype, roof, cpu, extra = run_web(args, wl)
    elif wl["kind"] == "wpr":
        value, ms, dtype, roof, cpu, extra = run_wpr(args, wl)
    elif wl["kind"] == "prfull":
        value, ms, dtype, roof, cpu, extra = run_prfull(args, wl)
    else:
        value, ms, dtype, roof, cpu, extra = run_pr(args, wl)
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": {"wpr": "synthetic (seeded R-MAT with seeded weights, generated on the device; seeded "
                        "personalization and loss vectors)",
                 "prfull": "synthetic (seeded R-MAT from the host generator, uploaded)"}.get(
            wl["kind"], "synthetic (seeded R-MAT + seeded update batches, generated on the device)"),
        "config": {"workload": wl["desc"], "scale": wl["scale"], "edge_factor": 16,
                   "batch_fraction": wl.get("frac"), "supernodes": supernodes_note(args, wl)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
