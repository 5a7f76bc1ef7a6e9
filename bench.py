#!/usr/bin/env python3
"""Streaming connected components on one MI355X: updates/s and p50/p99 query latency.

Workload (default): the 2^18-vertex serving loop.  A Chung-Lu stream (power-law exponent 2.1, mean
expected degree 64) of 400M updates, 15% of them deletes of present edges, arrives as batches of
log-uniform size in [1k, 1M]; after each batch come Poisson-many requests (on average one per
400k updates: 99% u-v point queries on uniform pairs, 1% full connected_components()).  The stream
is generated on the device before timing (inputs resident in HBM).

A step is one replay of the whole trace through the batched/served form on a fresh instance; W
untimed warmup steps precede K timed ones.  Beside it, on bounded prefixes of the same stream:
  * the plain GPU form (host gutters, one kernel launch per gutter flush, synchronous queries);
  * the CPU baseline: the oracle's restatement of the reference CPU path (per-vertex gutters,
    every host thread ingesting, per-thread delta sketches, OpenMP Borůvka).
Both forms' sketches are compared by digest on the same prefix.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "edge updates/sec + p50/p99 CC-query latency, 2^18-vertex stream, 1 GPU"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(REPO, "profiles", "k_update_tiles_chunglu_n262144_traffic.md")
VALU_FILE = os.path.join(REPO, "profiles", "k_update_tiles_chunglu_n262144_valu.md")
# wave-level VALU issue peak: 256 CUs x 4 SIMDs x one wave64 instruction per 4 cycles x 2.4 GHz
VALU_PEAK_GIPS = 256 * 2.4


def workload_key(args) -> str:
    return (f"n={args.n} updates={args.updates} del={args.del_frac} per_request={args.per_request} "
            f"seed={args.seed} sketch_seed={args.sketch_seed} gamma={args.gamma_bytes} tile={args.tile} "
            f"flush_threshold={args.flush_threshold}")


def measured_traffic(args):
    """HBM bytes per k_update_tiles launch from the committed rocprofv3 --pmc summary (FETCH_SIZE +
    WRITE_SIZE, each scaled by the same counter over a 1 GiB device copy), when it was measured on
    this exact workload; otherwise None."""
    try:
        with open(TRAFFIC_FILE) as f:
            kv = dict(line.split(":", 1) for line in f if line.startswith(("workload_key:", "traffic_bytes_per_launch:")))
    except OSError:
        return None
    if kv.get("workload_key", "").strip() != workload_key(args):
        return None
    try:
        return float(kv["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def measured_valu(args):
    """SQ_INSTS_VALU per k_update_tiles launch from the committed rocprofv3 --pmc summary, when it
    was measured on this exact workload; otherwise None."""
    try:
        with open(VALU_FILE) as f:
            kv = dict(line.split(":", 1) for line in f if line.startswith(("workload_key:", "valu_insts_per_launch:")))
    except OSError:
        return None
    if kv.get("workload_key", "").strip() != workload_key(args):
        return None
    try:
        return float(kv["valu_insts_per_launch"])
    except (KeyError, ValueError):
        return None


def pct(x, q):
    return float(np.percentile(x, q)) if len(x) else None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 18)
    ap.add_argument("--updates", type=int, default=400_000_000)
    ap.add_argument("--del-frac", type=float, default=0.15)
    ap.add_argument("--per-request", type=int, default=400_000, help="updates per request on average")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--sketch-seed", type=int, default=42)
    ap.add_argument("--gamma-bytes", type=int, default=4)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--flush-threshold", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=8_000_000, help="updates in the CPU-baseline prefix")
    ap.add_argument("--plain-sample", type=int, default=4_000_000, help="updates in the plain-form prefix")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, copy 1 GiB device-to-device 3 times (a plain copy of known "
                         "size that scales the FETCH_SIZE / WRITE_SIZE counters of a --pmc run)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}))
        return 2

    import torch

    import graphzeppelin_amd as gz
    from graphzeppelin_amd import gen

    if gz.device_count() < 1:
        raise SystemExit("bench.py needs a HIP device")
    n, N = args.n, args.updates
    t0 = time.time()
    log(f"generating Chung-Lu stream n={n} updates={N} del={args.del_frac} on the device")
    pairs, types, _ = gen.chung_lu_stream(n, N, args.del_frac, args.seed, device="cuda")
    del types
    torch.cuda.synchronize()
    log(f"stream ready in {time.time() - t0:.1f}s ({pairs.numel() * 4 / 2**30:.2f} GiB in HBM)")
    trace = gen.serving_trace(n, N, args.seed + 7, args.per_request)
    nq = int((trace["kind"] != 0).sum())
    log(f"trace: {trace.size} requests, {nq} queries")
    cfg = gz.CCAlgConfiguration(gamma_bytes=args.gamma_bytes, tile=args.tile, flush_threshold=args.flush_threshold)
    ptr = pairs.data_ptr()

    def one_step(profile: bool):
        alg = gz.CCSketchAlg(n, args.sketch_seed, cfg)
        if profile:
            alg.set_profiling(True)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        lat, ans, st = alg.serve_trace(ptr, True, trace)
        torch.cuda.synchronize()
        dt = time.perf_counter() - s0
        prof = alg.get_profile() if profile else None
        return alg, dt, lat, ans, st, prof

    for w in range(args.warmup):
        _, dt, _, _, st, _ = one_step(False)
        log(f"warmup {w}: {dt:.3f}s  {st['updates'] / dt / 1e6:.1f} M updates/s")
    times, lats, prof_tot, stats = [], [], None, []
    alg = None
    for k in range(args.steps):
        alg, dt, lat, ans, st, prof = one_step(True)
        times.append(dt)
        lats.append(lat)
        stats.append(st)
        if prof_tot is None:
            prof_tot = prof
        else:
            for key in ("update", "boruvka"):
                for f in ("launches", "total_ms", "alg_bytes", "items"):
                    prof_tot[key][f] += prof[key][f]
        log(f"step {k}: {dt:.3f}s  {st['updates'] / dt / 1e6:.1f} M updates/s  borůvka={st['boruvka_runs']} "
            f"cached={st['cached_answers']}")
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    qmask = trace["kind"] != 0
    qlat = np.concatenate([l[qmask] for l in lats]) * 1e3
    cc_lat = np.concatenate([l[trace["kind"] == 2] for l in lats]) * 1e3
    total_updates = sum(s["updates"] for s in stats)
    value = total_updates / sum(times)
    up = prof_tot["update"]
    achieved = up["alg_bytes"] / (up["total_ms"] / 1e3) / 1e9 if up["total_ms"] > 0 else 0.0
    br = prof_tot["boruvka"]
    br_gbs = br["alg_bytes"] / (br["total_ms"] / 1e3) / 1e9 if br["total_ms"] > 0 else 0.0

    def valu_roof(args, up):
        insts = measured_valu(args)
        if insts is None or up["launches"] == 0 or up["total_ms"] <= 0:
            return None
        gips = insts / (up["total_ms"] / up["launches"] / 1e3) / 1e9
        return {"insts_per_launch": insts, "achieved": gips, "peak": VALU_PEAK_GIPS,
                "unit": "G wave-instructions/s", "frac": gips / VALU_PEAK_GIPS}

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sum(times) / len(times),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": (f"serving loop: n={n} Chung-Lu(exponent 2.1, mean expected degree 64) stream of {N} "
                         f"updates ({args.del_frac:.0%} deletes of present edges) in log-uniform batches of "
                         f"1k-1M updates; Poisson requests, one per {args.per_request} updates on average "
                         f"(99% u-v point queries, 1% connected_components)"),
            "num_vertices": n,
            "updates": N,
            "requests": int(trace.size - (trace["kind"] == 0).sum()),
            "gamma_bytes": args.gamma_bytes,
            "sketch_seed": args.sketch_seed,
        },
        "p50_query_ms": pct(qlat, 50),
        "p99_query_ms": pct(qlat, 99),
        "p50_cc_query_ms": pct(cc_lat, 50),
        "p99_cc_query_ms": pct(cc_lat, 99),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_update_tiles (sketch update + delta write-back)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args),
            "launches": up["launches"],
            "avg_launch_ms": up["total_ms"] / max(1, up["launches"]),
            "alg_bytes_per_launch": up["alg_bytes"] / max(1, up["launches"]),
            "valu": valu_roof(args, up),
        },
        "boruvka_rounds": {
            "launches": br["launches"],
            "avg_round_ms": br["total_ms"] / max(1, br["launches"]),
            "achieved_GBs": br_gbs,
            "frac": br_gbs / HBM_PEAK_GBS,
        },
        "serve_stats": stats[-1],
    }

    # ---- plain GPU form and parity on a prefix ----
    if not args.no_plain:
        P = min(args.plain_sample, N)
        host = pairs[:P].cpu().numpy().astype(np.uint32)
        ptrace = gen.trace_prefix(trace, P)
        plain = gz.CCSketchAlg(n, args.sketch_seed, gz.CCAlgConfiguration(
            gamma_bytes=args.gamma_bytes, batched=False, tile=args.tile))
        log(f"plain form on the first {P} updates")
        plat, pans, pst = plain.plain_trace(host, ptrace)
        bat = gz.CCSketchAlg(n, args.sketch_seed, cfg)
        blat, bans, bst = bat.serve_trace(ptr, True, ptrace)
        pq = ptrace["kind"] != 0
        result["plain_gpu"] = {
            "value": pst["updates"] / pst["wall_s"],
            "unit": "updates/s",
            "sample": f"first {P} updates and their {int(pq.sum())} requests",
            "p50_query_ms": pct(plat[pq] * 1e3, 50),
            "p99_query_ms": pct(plat[pq] * 1e3, 99),
            "batched_same_sample_updates_per_s": bst["updates"] / bst["wall_s"],
            "speedup_batched_over_plain": (bst["updates"] / bst["wall_s"]) / (pst["updates"] / pst["wall_s"]),
            "digest_equal": plain.digest() == bat.digest(),
            "answers_equal": bool(np.array_equal(pans[pq], bans[pq])),
        }
        del plain, bat, host

    # ---- CPU baseline (the oracle's restatement of the reference CPU path) ----
    if not args.no_cpu:
        import oracle

        P = min(args.cpu_sample, N)
        host = pairs[:P].cpu().numpy().astype(np.uint32)
        ctrace = gen.trace_prefix(trace, P)
        q = ctrace[ctrace["kind"] != 0]
        qpos = []
        upto = 0
        for r in ctrace:
            if r["kind"] == 0:
                upto = int(r["offset"] + r["count"])
            else:
                qpos.append(upto)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"CPU baseline on the first {P} updates, {q.size} queries, {threads} threads")
        ref = oracle.CCAlg(n, args.sketch_seed, gamma_bytes=args.gamma_bytes)
        wall, clat, cans = ref.baseline_trace(host, np.array(qpos, dtype=np.uint64),
                                              (q["kind"] == 2).astype(np.int32), q["u"], q["v"], threads)
        bat = gz.CCSketchAlg(n, args.sketch_seed, cfg)
        _, bans, _ = bat.serve_trace(ptr, True, ctrace)
        img = ref.dump()
        cpu_digest = gz.digest_of_image(img, bat.sketch_bytes() // (8 + args.gamma_bytes), args.gamma_bytes)
        result["cpu_baseline"] = {
            "value": P / wall,
            "unit": "updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"first {P} updates and their {q.size} requests (oracle restatement of the reference "
                      f"CPU path, {threads} host threads)",
            "p50_query_ms": pct(clat * 1e3, 50),
            "p99_query_ms": pct(clat * 1e3, 99),
            "digest_equal": cpu_digest == bat.digest(),
            "answers_equal": bool(np.array_equal(cans, bans[ctrace["kind"] != 0])),
        }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
