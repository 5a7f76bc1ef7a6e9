"""Multi-tenant serving at its stated scale on the GPU: 32 graph streams with n log-uniform in
[2^12, 2^20], each an R-MAT(0.57, 0.19, 0.19, 0.05) stream of 8 n log2(n) updates with 10% deletes
of present edges (drawn in chunks from torch's counter-based generator), admitted by sketch
footprint against the device's free memory and replayed as one mixed trace: every tenant's stream
arrives at a steady rate, with connected_components() and u-v point queries at random positions and
one connected_components() at the end of each stream.

* every answer of every tenant equals the exact answer over that tenant's edges live at that point;
* every tenant's final partition equals the exact components of its final edge set;
* tenants with n <= 2^15: the final sketches' digest equals that of the oracle's image (bit-exact);
* alone and in the mix: four tenants replayed alone through the single-stream loop give the same
  answers and sketch digests as in the mix.
"""
import time

import numpy as np
import pytest

import exact_cc
import graphzeppelin_amd as gz
import oracle

pytestmark = pytest.mark.gpu

T, SEED, RATE = 32, 5, 1.5e9  # tenants, workload seed, offered updates/s of the whole mix


def say(m):
    print(m, flush=True)


def test_config5_multi_tenant_exact():
    import torch

    from graphzeppelin_amd import gen

    t0 = time.time()
    tenants = gen.multi_tenant_workload(T, SEED)
    trace = gen.multi_tenant_trace(tenants, SEED + 1, RATE)
    torch.cuda.synchronize()
    ns = [n for n, _ in tenants]
    total = sum(int(p.shape[0]) for _, p in tenants)
    say(f"config5: {T} tenants, n {min(ns)}..{max(ns)}, {total} updates, {trace.size} requests "
        f"(generated in {time.time() - t0:.0f}s)")
    seeds = [900 + t for t in range(T)]
    cfg = gz.CCAlgConfiguration()
    need = sum(gz.footprint_bytes(n, cfg) for n in ns)
    srv = gz.Server(0, workers=4)
    say(f"config5: footprints {need / 2**30:.1f} GiB of a {srv.budget_bytes / 2**30:.1f} GiB budget")
    t0 = time.perf_counter()
    lat, ans, dig, st, parents = srv.run_trace(trace, [p.data_ptr() for _, p in tenants], ns, seeds, cfg,
                                               want_parents=True)
    wall = time.perf_counter() - t0
    q = trace["kind"] != gz.REQ_UPDATES
    say(f"config5: mix replayed in {wall:.2f}s ({st['updates'] / st['wall_s'] / 1e6:.0f} M updates/s), "
        f"{st['queries']} queries, p50/p99 latency {np.percentile(lat[q], 50) * 1e3:.1f}/"
        f"{np.percentile(lat[q], 99) * 1e3:.1f} ms")
    assert st["updates"] == total and st["queries"] == int(q.sum())
    assert srv.reserved_bytes == 0
    checked = 0
    for t, (n, pairs) in enumerate(tenants):
        sel = np.flatnonzero(trace["tenant"] == t)
        life = exact_cc.EdgeLifetimes(pairs)
        upto, cache = 0, {}
        for i in sel:
            r = trace[i]
            if r["kind"] == gz.REQ_UPDATES:
                upto = int(r["offset"] + r["count"])
                continue
            if upto not in cache:
                c, lab = life.components(n, upto)
                cache[upto] = (c, lab.cpu().numpy())
            c, lab = cache[upto]
            want = c if r["kind"] == gz.REQ_CC else int(lab[r["u"]] == lab[r["v"]])
            assert ans[i] == want, f"tenant {t} (n={n}) request {i} at {upto}: got {ans[i]} want {want}"
            checked += 1
        c, lab = life.components(n, int(pairs.shape[0]))
        assert np.array_equal(oracle.canonical(parents[t]), lab.cpu().numpy().astype(np.uint32)), f"tenant {t}"
        del life, lab, cache
        if n <= 1 << 15:
            ref = oracle.CCAlg(n, seeds[t])
            ref.ingest_pairs(pairs.cpu().numpy().astype(np.uint32), with_pre_insert=False)
            assert int(dig[t]) == gz.digest_of_image(ref.dump(), ref.sketch_bytes() // 12, 4), f"tenant {t}"
            del ref
        say(f"config5: tenant {t} n={n}: {len(sel)} requests exact")
    torch.cuda.empty_cache()
    # alone and in the mix: the smallest, the median and the two largest tenants
    order = np.argsort(ns)
    for t in [int(order[0]), int(order[T // 2]), int(order[-2]), int(order[-1])]:
        n, pairs = tenants[t]
        sel = trace["tenant"] == t
        alone = gz.CCSketchAlg(n, seeds[t], cfg)
        _, aans, _ = alone.serve_trace(pairs.data_ptr(), True, trace[sel].copy())
        sq = trace[sel]["kind"] != gz.REQ_UPDATES
        assert np.array_equal(ans[sel][sq], aans[sq]), f"tenant {t}"
        assert int(dig[t]) == alone.digest(), f"tenant {t}"
        del alone
    say(f"config5: {checked} answers exact; alone == mix for 4 tenants")
