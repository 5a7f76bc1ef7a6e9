#!/usr/bin/env python3
"""Streaming connected components on one MI355X: updates/s and p50/p99 query latency.

Workload (default): the 2^18-vertex serving loop.  A Chung-Lu stream (power-law exponent 2.1, mean
expected degree 64) of 400M updates, 15% of them deletes of present edges, arrives as batches of
log-uniform size in [1k, 1M]; after each batch come Poisson-many requests (on average one per
400k updates: 99% u-v point queries on uniform pairs, 1% full connected_components()).  The stream
is generated on the device before timing (inputs resident in HBM).

A step is one replay of the whole trace through the batched/served form on a fresh instance; W
untimed warmup steps precede K timed ones, which run without kernel timers.  One more replay with
HIP-event kernel timers gives the roofline figures.  Beside it (each leg can be switched off):
  * query-rate sensitivity: the same stream with one request per 100k and per 1.6M updates;
  * 8-byte checksums (16-byte buckets, the wider reading of the reference's vec_hash_t);
  * host-fed: the same stream in pinned and in pageable host memory (PCIe inside the timed region);
  * multi-tenant: 32 R-MAT streams with n log-uniform in [2^12, 2^20] served as one mix, saturating
    and at half of the saturating rate;
  * the plain GPU form (host gutters, one kernel launch per gutter flush, synchronous queries);
  * the CPU baseline: the oracle's restatement of the reference CPU path (per-vertex gutters,
    every host thread ingesting, per-thread delta sketches, OpenMP Borůvka).
The plain and CPU forms are compared with the batched form by sketch digest and answers on the same
prefix.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "edge updates/sec + p50/p99 CC-query latency, 2^18-vertex stream, 1 GPU"
HBM_PEAK_GBS = 8000.0
# wave-level VALU issue peak: 256 CUs x 4 SIMDs x one wave64 instruction per 4 cycles x 2.4 GHz
VALU_PEAK_GIPS = 256 * 2.4
TRAFFIC_FILE = os.path.join(REPO, "profiles", "k_update_tiles_chunglu_n262144_traffic.md")
VALU_FILE = os.path.join(REPO, "profiles", "k_update_tiles_chunglu_n262144_valu.md")
ALL_LEGS = ("sensitivity", "gamma8", "host", "multitenant")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def source_hash() -> str:
    """Hash of the library's sources: a committed counter summary counts only for the build it was
    measured on."""
    d = os.path.join(REPO, "graphzeppelin_amd", "csrc")
    h = hashlib.sha256()
    for f in sorted(os.listdir(d)):
        if f.endswith((".hip", ".cpp", ".hpp")):
            h.update(f.encode())
            with open(os.path.join(d, f), "rb") as fh:
                h.update(fh.read())
    return h.hexdigest()[:16]


def workload_key(args) -> str:
    return (f"n={args.n} updates={args.updates} del={args.del_frac} per_request={args.per_request} "
            f"seed={args.seed} sketch_seed={args.sketch_seed} gamma={args.gamma_bytes} tile={args.tile} "
            f"flush_threshold={args.flush_threshold}")


def committed_figure(path: str, field: str, args):
    """A per-launch counter figure from a committed rocprofv3 --pmc summary, when it was measured on
    this exact workload and this exact source tree; otherwise None."""
    try:
        with open(path) as f:
            kv = {k.strip(): v.strip() for k, v in (line.split(":", 1) for line in f if ":" in line)}
    except OSError:
        return None
    if kv.get("workload_key") != workload_key(args) or kv.get("source_hash") != source_hash():
        return None
    try:
        return float(kv[field])
    except (KeyError, ValueError):
        return None


def pct(x, q):
    return float(np.percentile(x, q)) if len(x) else None


def host_cpu() -> dict:
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "cpus_available": len(os.sched_getaffinity(0))}


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
    ap.add_argument("--legs", default=",".join(ALL_LEGS),
                    help=f"comma-separated side legs to run, from {ALL_LEGS}, or 'none'")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, copy 1 GiB device-to-device 3 times (a plain copy of known "
                         "size that scales the FETCH_SIZE / WRITE_SIZE counters of a --pmc run)")
    args = ap.parse_args()
    legs = set() if args.legs in ("", "none") else set(args.legs.split(","))
    if legs - set(ALL_LEGS):
        ap.error(f"unknown legs {sorted(legs - set(ALL_LEGS))}")

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
    qmask = trace["kind"] != gz.REQ_UPDATES
    log(f"trace: {trace.size} requests, {int(qmask.sum())} queries")
    cfg = gz.CCAlgConfiguration(gamma_bytes=args.gamma_bytes, tile=args.tile, flush_threshold=args.flush_threshold)
    ptr = pairs.data_ptr()

    def replay(tr, config, src_ptr=ptr, on_device=True, profile=False):
        alg = gz.CCSketchAlg(n, args.sketch_seed, config)
        if profile:
            alg.set_profiling(True)
        torch.cuda.synchronize()
        s0 = time.perf_counter()
        lat, ans, st = alg.serve_trace(src_ptr, on_device, tr)
        torch.cuda.synchronize()
        dt = time.perf_counter() - s0
        prof = alg.get_profile() if profile else None
        return dt, lat, ans, st, prof

    def summary(runs, tr):
        q = tr["kind"] != gz.REQ_UPDATES
        ql = np.concatenate([r[1][q] for r in runs]) * 1e3
        upd = sum(r[3]["updates"] for r in runs)
        return {"value": upd / sum(r[0] for r in runs), "unit": "updates/s", "steps": len(runs),
                "ms_per_step": 1e3 * sum(r[0] for r in runs) / len(runs), "queries": int(q.sum()),
                "p50_query_ms": pct(ql, 50), "p99_query_ms": pct(ql, 99),
                "boruvka_runs": runs[-1][3]["boruvka_runs"]}

    # ---- the timed steps (no kernel timers) ----
    for w in range(args.warmup):
        dt, _, _, st, _ = replay(trace, cfg)
        log(f"warmup {w}: {dt:.3f}s  {st['updates'] / dt / 1e6:.1f} M updates/s")
    runs = []
    for k in range(args.steps):
        r = replay(trace, cfg)
        runs.append(r)
        log(f"step {k}: {r[0]:.3f}s  {r[3]['updates'] / r[0] / 1e6:.1f} M updates/s  borůvka={r[3]['boruvka_runs']} "
            f"cached={r[3]['cached_answers']}")
    times = [r[0] for r in runs]
    value = sum(r[3]["updates"] for r in runs) / sum(times)
    qlat = np.concatenate([r[1][qmask] for r in runs]) * 1e3
    cc_lat = np.concatenate([r[1][trace["kind"] == gz.REQ_CC] for r in runs]) * 1e3

    # ---- one more replay with HIP-event kernel timers, for the roofline ----
    pdt, _, _, _, prof = replay(trace, cfg, profile=True)
    log(f"profiled replay: {pdt:.3f}s")
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    up, br = prof["update"], prof["boruvka"]
    avg_ms = up["total_ms"] / max(1, up["launches"])
    achieved = up["alg_bytes"] / (up["total_ms"] / 1e3) / 1e9 if up["total_ms"] > 0 else 0.0
    br_gbs = br["alg_bytes"] / (br["total_ms"] / 1e3) / 1e9 if br["total_ms"] > 0 else 0.0
    insts = committed_figure(VALU_FILE, "valu_insts_per_launch", args)
    valu = None
    if insts is not None and avg_ms > 0:
        gips = insts / (avg_ms / 1e3) / 1e9
        valu = {"insts_per_launch": insts, "achieved": gips, "peak": VALU_PEAK_GIPS,
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
                         f"(99% u-v point queries, 1% connected_components); stream resident in HBM"),
            "num_vertices": n,
            "updates": N,
            "requests": int(qmask.sum()),
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
            "traffic": committed_figure(TRAFFIC_FILE, "traffic_bytes_per_launch", args),
            "launches": up["launches"],
            "avg_launch_ms": avg_ms,
            "alg_bytes_per_launch": up["alg_bytes"] / max(1, up["launches"]),
            "timed_in": "a separate replay with HIP-event kernel timers (the timed steps run without them)",
            "profiled_replay_ms": 1e3 * pdt,
            "valu": valu,
        },
        "boruvka_rounds": {
            "queries": br["launches"],
            "rounds": br["items"],
            "avg_query_ms": br["total_ms"] / max(1, br["launches"]),
            "achieved_GBs": br_gbs,
            "frac": br_gbs / HBM_PEAK_GBS,
            "bytes_rule": "per round: (D+1)*B + 8 bytes for every vertex the round regroups",
        },
        "serve_stats": runs[-1][3],
        "source_hash": source_hash(),
    }

    # ---- side legs ----
    if "sensitivity" in legs:
        out = []
        for per in (100_000, 1_600_000):
            tr = gen.serving_trace(n, N, args.seed + 7, per)
            rs = [replay(tr, cfg) for _ in range(2)]
            s = summary(rs, tr)
            s["per_request"] = per
            out.append(s)
            log(f"sensitivity: one request per {per} updates: {s['value'] / 1e6:.1f} M updates/s, "
                f"{s['queries']} queries, p50 {s['p50_query_ms']:.2f} ms")
        result["query_rate_sensitivity"] = out
    if "gamma8" in legs and args.gamma_bytes != 8:
        c8 = gz.CCAlgConfiguration(gamma_bytes=8, tile=args.tile, flush_threshold=args.flush_threshold)
        replay(trace, c8)
        rs = [replay(trace, c8) for _ in range(2)]
        result["gamma8"] = summary(rs, trace)
        log(f"gamma 8: {result['gamma8']['value'] / 1e6:.1f} M updates/s")
    if "host" in legs:
        host_legs = {}
        pinned = torch.empty(pairs.shape, dtype=pairs.dtype, pin_memory=True)
        pinned.copy_(pairs)
        rs = [replay(trace, cfg, pinned.data_ptr(), False) for _ in range(2)]
        host_legs["pinned"] = summary(rs, trace)
        del pinned
        pageable = pairs.cpu()
        rs = [replay(trace, cfg, pageable.data_ptr(), False) for _ in range(2)]
        host_legs["pageable"] = summary(rs, trace)
        del pageable
        host_legs["note"] = "the stream in host memory: every update crosses PCIe inside the timed region"
        result["host_fed"] = host_legs
        log(f"host-fed: pinned {host_legs['pinned']['value'] / 1e6:.1f}, pageable "
            f"{host_legs['pageable']['value'] / 1e6:.1f} M updates/s")
    if "multitenant" in legs:
        result["multi_tenant"] = multi_tenant_leg(gz, gen, torch)

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
        pq = ptrace["kind"] != gz.REQ_UPDATES
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
        cq = ctrace["kind"] != gz.REQ_UPDATES
        q = ctrace[cq]
        qpos = []
        upto = 0
        for r in ctrace:
            if r["kind"] == gz.REQ_UPDATES:
                upto = int(r["offset"] + r["count"])
            else:
                qpos.append(upto)
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"CPU baseline on the first {P} updates, {q.size} queries, {threads} threads")
        ref = oracle.CCAlg(n, args.sketch_seed, gamma_bytes=args.gamma_bytes)
        wall, clat, cans = ref.baseline_trace(host, np.array(qpos, dtype=np.uint64),
                                              (q["kind"] == gz.REQ_CC).astype(np.int32), q["u"], q["v"], threads)
        bat = gz.CCSketchAlg(n, args.sketch_seed, cfg)
        torch.cuda.synchronize()
        b0 = time.perf_counter()
        blat, bans, bst = bat.serve_trace(ptr, True, ctrace)
        torch.cuda.synchronize()
        bdt = time.perf_counter() - b0
        img = ref.dump()
        cpu_digest = gz.digest_of_image(img, bat.sketch_bytes() // (8 + args.gamma_bytes), args.gamma_bytes)
        result["cpu_baseline"] = {
            "value": P / wall,
            "unit": "updates/s",
            "cores": threads,
            "kind": "port",
            "host_cpu": host_cpu(),
            "sample": f"first {P} updates and their {q.size} requests (oracle restatement of the reference "
                      f"CPU path, {threads} host threads)",
            "p50_query_ms": pct(clat * 1e3, 50),
            "p99_query_ms": pct(clat * 1e3, 99),
            "gpu_same_sample": {"value": P / bdt, "unit": "updates/s", "p50_query_ms": pct(blat[cq] * 1e3, 50),
                                "p99_query_ms": pct(blat[cq] * 1e3, 99)},
            "digest_equal": cpu_digest == bat.digest(),
            "answers_equal": bool(np.array_equal(cans, bans[cq])),
        }
    print(json.dumps(result), flush=True)
    return 0


def multi_tenant_leg(gz, gen, torch) -> dict:
    """32 R-MAT streams (n log-uniform in [2^12, 2^20], 8 n log2 n updates, 10% deletes) served as
    one mix: everything arriving at once (aggregate throughput), then every stream arriving at a
    steady rate, together a quarter and half of that throughput (latency under load)."""
    T, seed = 32, 5
    t0 = time.time()
    tenants = gen.multi_tenant_workload(T, seed)
    torch.cuda.synchronize()
    ns = [nv for nv, _ in tenants]
    total = sum(int(p.shape[0]) for _, p in tenants)
    log(f"multi-tenant: {T} streams, {total} updates, generated in {time.time() - t0:.1f}s")
    seeds = [900 + t for t in range(T)]
    ptrs = [p.data_ptr() for _, p in tenants]
    cfg = gz.CCAlgConfiguration()
    out = {"tenants": T, "updates": total, "n_min": min(ns), "n_max": max(ns),
           "footprint_GiB": sum(gz.footprint_bytes(nv, cfg) for nv in ns) / 2**30}

    def run(rate):
        tr = gen.multi_tenant_trace(tenants, seed + 1, rate)
        srv = gz.Server(0, workers=4)
        lat, ans, dig, st = srv.run_trace(tr, ptrs, ns, seeds, cfg)
        q = tr["kind"] != gz.REQ_UPDATES
        return {"value": st["updates"] / st["wall_s"], "unit": "updates/s", "wall_s": st["wall_s"],
                "queries": int(q.sum()), "p50_query_ms": pct(lat[q] * 1e3, 50), "p99_query_ms": pct(lat[q] * 1e3, 99),
                "offered_updates_per_s": rate}

    sat = run(None)
    out["saturating"] = sat
    # steady arrivals make smaller flushes than everything arriving at once, so a fraction of the
    # saturating rate need not be sustainable: each load point reports what it achieved
    out["quarter_load"] = run(0.25 * sat["value"])
    out["half_load"] = run(0.5 * sat["value"])
    log(f"multi-tenant: {sat['value'] / 1e6:.1f} M updates/s saturating; p50/p99 at a quarter of that "
        f"{out['quarter_load']['p50_query_ms']:.1f}/{out['quarter_load']['p99_query_ms']:.1f} ms, at half "
        f"{out['half_load']['p50_query_ms']:.1f}/{out['half_load']['p99_query_ms']:.1f} ms")
    del tenants
    torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    sys.exit(main())
