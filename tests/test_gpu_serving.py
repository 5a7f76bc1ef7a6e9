"""Serving loops on the GPU: the batched/served form, the plain form and the multi-tenant server.

Every answer is checked against an exact DSU over the stream prefix it must reflect, forms are
compared by sketch digest, and each tenant's results in a mixed, admission-limited run must equal
its results alone.
"""
import numpy as np
import pytest

import graphzeppelin_amd as gz
import oracle
from graphzeppelin_amd import streams

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)


def trace_for(n, total, seed, per_req, bmin=500, bmax=20000):
    rng = np.random.default_rng(seed)
    rows, off = [], 0
    while off < total:
        b = min(int(np.exp(rng.uniform(np.log(bmin), np.log(bmax)))), total - off)
        rows.append((gz.REQ_UPDATES, 0, off, b, 0, 0, 0.0))
        off += b
        for _ in range(rng.poisson(b / per_req)):
            if rng.random() < 0.2:
                rows.append((gz.REQ_CC, 0, 0, 0, 0, 0, 0.0))
            else:
                u, v = rng.integers(0, n, 2)
                rows.append((gz.REQ_POINT, 0, 0, 0, int(u), int(v), 0.0))
    return np.array(rows, dtype=gz.REQUEST_DTYPE)


def expected_answers(n, pairs, trace):
    ans = np.zeros(trace.size, dtype=np.uint32)
    upto = 0
    cache = {}
    for i, r in enumerate(trace):
        if r["kind"] == gz.REQ_UPDATES:
            upto = int(r["offset"] + r["count"])
            continue
        if upto not in cache:
            cache[upto] = oracle.exact_cc(n, pairs[:upto])
        c, lab = cache[upto]
        ans[i] = c if r["kind"] == gz.REQ_CC else int(lab[r["u"]] == lab[r["v"]])
    return ans


@pytest.mark.parametrize("eager", [True, False])
def test_serve_trace_answers_exact(eager):
    import torch

    say(f"serve_trace n=4096 eager={eager}")
    n = 4096
    pairs, types = streams.dynamic_erdos_stream(n, 0.002, 0.3, 0.1, 2, 17)
    trace = trace_for(n, pairs.shape[0], 3, 2000)
    dev = torch.from_numpy(pairs.astype(np.int32)).cuda()
    alg = gz.CCSketchAlg(n, 99, gz.CCAlgConfiguration(eager_dsu=eager, flush_threshold=1 << 14))
    lat, ans, st = alg.serve_trace(dev.data_ptr(), True, trace)
    q = trace["kind"] != 0
    assert st["updates"] == pairs.shape[0] and st["queries"] == int(q.sum())
    assert np.array_equal(ans[q], expected_answers(n, pairs, trace)[q])
    assert np.all(lat[q] > 0)
    ref = gz.CCSketchAlg(n, 99)
    ref.ingest_pairs(pairs)
    assert alg.digest() == ref.digest()
    img = ref.dump_sketches()
    assert gz.digest_of_image(img, ref.sketch_bytes() // 12, 4) == ref.digest()


@pytest.mark.parametrize("eager", [True, False])
def test_out_of_range_device_pair_is_reported_by_a_query(eager):
    """A device pair naming a vertex >= n is skipped and reported as INVALID_ARG by the next query,
    also when that query runs Borůvka (the query's DSU reset keeps the flag); the instance then
    answers from the valid pairs alone."""
    import torch

    say(f"out-of-range device pair eager={eager}")
    n = 512
    pairs, _ = streams.erdos_renyi_churn_stream(n, 0.02, 5)
    bad = pairs.copy()
    bad[len(bad) // 2] = (3, n)  # one pair names vertex n
    dev = torch.from_numpy(bad.astype(np.int32)).cuda()
    alg = gz.CCSketchAlg(n, 4, gz.CCAlgConfiguration(eager_dsu=eager))
    trace = np.array([(gz.REQ_UPDATES, 0, 0, bad.shape[0], 0, 0, 0.0), (gz.REQ_POINT, 0, 0, 0, 1, 2, 0.0)],
                     dtype=gz.REQUEST_DTYPE)
    with pytest.raises(gz.GZError) as ei:
        alg.serve_trace(dev.data_ptr(), True, trace)
    assert ei.value.code == 1, ei.value  # GZ_ERR_INVALID_ARG
    # the flag was cleared where it was reported; the sketches hold every valid pair
    keep = np.ones(bad.shape[0], dtype=bool)
    keep[len(bad) // 2] = False
    c, lab = oracle.exact_cc(n, bad[keep])
    cc = alg.connected_components()
    assert cc.size() == c and np.array_equal(cc.canonical_labels(), lab)
    ref = oracle.CCAlg(n, 4)
    ref.ingest_pairs(bad[keep])
    assert np.array_equal(alg.dump_sketches(), ref.dump())


def test_cached_first_query_on_recycled_memory():
    """A query the eager DSU still answers keeps that DSU and skips every Borůvka kernel.  On a new
    instance whose device memory was just used and freed by another instance (so it is not zero),
    the round kernels of such a first query must still only read in-range state; answers exact."""
    import torch

    say("cached first query on recycled memory")
    n = 4096
    pairs, _ = streams.dynamic_erdos_stream(n, 0.002, 0.3, 0.1, 2, 31)
    dev = torch.from_numpy(pairs.astype(np.int32)).cuda()
    dirty = gz.CCSketchAlg(n, 5, gz.CCAlgConfiguration(eager_dsu=False))
    dirty.serve_trace(dev.data_ptr(), True, trace_for(n, pairs.shape[0], 9, 3000))
    dirty.close()
    # insert-only prefix: the eager DSU is valid at the first query, which is answered from it
    ins = pairs[:2000]
    trace = np.array([(gz.REQ_UPDATES, 0, 0, ins.shape[0], 0, 0, 0.0), (gz.REQ_CC, 0, 0, 0, 0, 0, 0.0),
                      (gz.REQ_POINT, 0, 0, 0, int(ins[0, 0]), int(ins[0, 1]), 0.0),
                      (gz.REQ_UPDATES, 0, ins.shape[0], pairs.shape[0] - ins.shape[0], 0, 0, 0.0),
                      (gz.REQ_CC, 0, 0, 0, 0, 0, 0.0)], dtype=gz.REQUEST_DTYPE)
    alg = gz.CCSketchAlg(n, 5)
    _, ans, st = alg.serve_trace(dev.data_ptr(), True, trace)
    assert st["cached_answers"] >= 1
    c1, _ = oracle.exact_cc(n, ins)
    c2, _ = oracle.exact_cc(n, pairs)
    assert list(ans[[1, 2, 4]]) == [c1, 1, c2]
    ref = oracle.CCAlg(n, 5)
    ref.ingest_pairs(pairs)
    assert np.array_equal(alg.dump_sketches(), ref.dump())


def test_query_kept_cached_right_after_borůvka_keeps_forest_checks():
    """Two query calls with no update between: the first runs Borůvka (its reset empties the
    spanning-forest set, which the next eager batch refills), the second keeps the fresh DSU.  The
    second must leave the refill pending, or later deletions of forest edges would go unnoticed:
    deleting every live edge afterwards must give n components."""
    import torch

    say("cached query right after Borůvka")
    n = 2048
    pairs, _ = streams.dynamic_erdos_stream(n, 0.003, 0.3, 0.1, 2, 41)
    live = streams.final_edge_set(pairs)
    alg = gz.CCSketchAlg(n, 8)
    dev = torch.from_numpy(pairs.astype(np.int32)).cuda()
    t1 = np.array([(gz.REQ_UPDATES, 0, 0, pairs.shape[0], 0, 0, 0.0), (gz.REQ_CC, 0, 0, 0, 0, 0, 0.0)],
                  dtype=gz.REQUEST_DTYPE)
    _, a1, s1 = alg.serve_trace(dev.data_ptr(), True, t1)
    c, _ = oracle.exact_cc(n, pairs)
    assert a1[1] == c and s1["boruvka_runs"] == 1  # the stream's deletions invalidated the eager DSU
    dels = torch.from_numpy(live.astype(np.int32)).cuda()
    t2 = np.array([(gz.REQ_CC, 0, 0, 0, 0, 0, 0.0), (gz.REQ_UPDATES, 0, 0, live.shape[0], 0, 0, 0.0),
                   (gz.REQ_CC, 0, 0, 0, 0, 0, 0.0)], dtype=gz.REQUEST_DTYPE)
    _, a2, s2 = alg.serve_trace(dels.data_ptr(), True, t2)
    assert a2[0] == c and s2["cached_answers"] >= 1  # kept the DSU Borůvka had just built
    assert a2[2] == n  # every edge deleted


def test_plain_and_batched_forms_agree():
    say("plain vs batched n=2048")
    n = 2048
    pairs, _ = streams.dynamic_erdos_stream(n, 0.004, 0.4, 0.1, 2, 23)
    trace = trace_for(n, pairs.shape[0], 5, 3000)
    plain = gz.CCSketchAlg(n, 7, gz.CCAlgConfiguration(batched=False))
    plat, pans, pst = plain.plain_trace(pairs, trace)
    bat = gz.CCSketchAlg(n, 7)
    blat, bans, bst = bat.serve_trace(pairs.ctypes.data, False, trace)
    q = trace["kind"] != 0
    assert np.array_equal(pans[q], bans[q])
    assert np.array_equal(bans[q], expected_answers(n, pairs, trace)[q])
    assert plain.digest() == bat.digest()
    ref = oracle.CCAlg(n, 7)
    ref.ingest_pairs(pairs)
    assert gz.digest_of_image(ref.dump(), ref.sketch_bytes() // 12, 4) == bat.digest()


def _tenant_streams(T, seed):
    import torch

    from graphzeppelin_amd import gen

    rng = np.random.default_rng(seed)
    out = []
    for t in range(T):
        scale = int(rng.integers(10, 14))
        n = 1 << scale
        m = 8 * n * scale
        p, ty, fin = gen.rmat_stream(scale, m, 0.1, 1000 + t, device="cuda")
        out.append((n, p.contiguous()))
    return out


def _mixed_trace(tenants, seed, per_req=20000):
    rng = np.random.default_rng(seed)
    per = []
    for t, (n, p) in enumerate(tenants):
        rows, off, total = [], 0, p.shape[0]
        while off < total:
            b = min(int(np.exp(rng.uniform(np.log(1000), np.log(100000)))), total - off)
            rows.append((gz.REQ_UPDATES, t, off, b, 0, 0, 0.0))
            off += b
            for _ in range(rng.poisson(b / per_req)):
                if rng.random() < 0.1:
                    rows.append((gz.REQ_CC, t, 0, 0, 0, 0, 0.0))
                else:
                    u, v = rng.integers(0, n, 2)
                    rows.append((gz.REQ_POINT, t, 0, 0, int(u), int(v), 0.0))
        per.append(rows)
    # interleave tenants round-robin, arrival times spread over ~0.2 s
    mixed, idx = [], [0] * len(per)
    while any(idx[t] < len(per[t]) for t in range(len(per))):
        for t in range(len(per)):
            if idx[t] < len(per[t]):
                mixed.append(per[t][idx[t]])
                idx[t] += 1
    arr = np.array(mixed, dtype=gz.REQUEST_DTYPE)
    arr["arrival_s"] = np.linspace(0, 0.2, arr.size)
    return arr


def _exact_answers(n, pairs_dev, trace):
    import exact_cc

    life = exact_cc.EdgeLifetimes(pairs_dev)
    out = np.zeros(trace.size, dtype=np.uint32)
    upto = 0
    for i, r in enumerate(trace):
        if r["kind"] == gz.REQ_UPDATES:
            upto = int(r["offset"] + r["count"])
            continue
        c, lab = life.components(n, upto)
        out[i] = c if r["kind"] == gz.REQ_CC else int(lab[r["u"]] == lab[r["v"]])
    return out


def test_multi_tenant_mix_equals_alone():
    say("multi-tenant mix")
    tenants = _tenant_streams(6, 11)
    trace = _mixed_trace(tenants, 12)
    ns = [n for n, _ in tenants]
    seeds = [500 + t for t in range(len(tenants))]
    cfg = gz.CCAlgConfiguration(flush_threshold=1 << 20)
    fps = [gz.footprint_bytes(n, cfg) for n in ns]
    budget = int(sum(sorted(fps)[-3:]) * 1.01)  # at most about three tenants resident at once
    srv = gz.Server(budget, workers=3)
    lat, ans, dig, st = srv.run_trace(trace, [p.data_ptr() for _, p in tenants], ns, seeds, cfg)
    assert srv.reserved_bytes == 0
    q = trace["kind"] != 0
    assert st["queries"] == int(q.sum())
    for t, (n, p) in enumerate(tenants):
        sel = trace["tenant"] == t
        sub = trace[sel].copy()
        alone = gz.CCSketchAlg(n, seeds[t], cfg)
        alat, aans, ast = alone.serve_trace(p.data_ptr(), True, sub)
        sq = sub["kind"] != 0
        assert np.array_equal(aans[sq], _exact_answers(n, p, sub)[sq]), f"tenant {t} alone"
        assert np.array_equal(ans[sel][sq], aans[sq]), f"tenant {t}"
        assert int(dig[t]) == alone.digest(), f"tenant {t}"
        assert np.all(lat[sel][sq] >= 0)


def test_admission_rejects_oversized_tenant():
    small = gz.footprint_bytes(64)
    srv = gz.Server(small + 1, workers=1)
    with pytest.raises(gz.AdmissionError):
        srv.admit(1 << 16, 1)
    t = srv.admit(64, 1)
    with pytest.raises(gz.AdmissionError):  # the budget is now taken
        srv.admit(64, 2)
    assert srv.reserved_bytes > 0
    srv.release(t)
    assert srv.reserved_bytes == 0
