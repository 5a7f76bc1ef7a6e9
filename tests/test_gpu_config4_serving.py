"""The serving workload at its full size on the GPU: n = 2^18, a Chung-Lu stream (exponent 2.1,
mean expected degree 64) of 400M updates with 15% deletes of present edges, sketch seed 42, and
the request trace the benchmark replays (log-uniform batches of 1k-1M updates, one request per
400k updates on average, 99% u-v point queries and 1% connected_components()).

* Bit-exact: after the trace's first 20M updates, served through the serving loop with its
  queries, every vertex's packed bucket array equals the oracle's byte for byte (2.8 GB).
* Answers: the whole trace is replayed through the serving loop (as the benchmark times it) and
  every answer of every query group equals the exact answer over the edges live at that point of
  the stream; afterwards the partition equals the exact components of the final edge set.

The exact answers come from the stream's edge lifetimes and min-label propagation in torch
(tests/exact_cc.py), independent of the library's kernels.
"""
import time

import numpy as np
import pytest

import exact_cc
import graphzeppelin_amd as gz
import oracle

pytestmark = pytest.mark.gpu

N_VERT, N_UPD, DEL_FRAC, STREAM_SEED, SKETCH_SEED = 1 << 18, 400_000_000, 0.15, 1, 42
PREFIX = 20_000_000


def say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def workload():
    import torch

    from graphzeppelin_amd import gen

    t0 = time.time()
    pairs, _, _ = gen.chung_lu_stream(N_VERT, N_UPD, DEL_FRAC, STREAM_SEED, device="cuda")
    trace = gen.serving_trace(N_VERT, N_UPD, STREAM_SEED + 7, 400_000)
    torch.cuda.synchronize()
    say(f"config4: stream and trace ready in {time.time() - t0:.1f}s, {trace.size} requests")
    yield pairs, trace
    del pairs
    torch.cuda.empty_cache()


def test_config4_prefix_sketches_bit_exact(workload):
    from graphzeppelin_amd import gen

    pairs, trace = workload
    ptrace = gen.trace_prefix(trace, PREFIX)
    alg = gz.CCSketchAlg(N_VERT, SKETCH_SEED)
    _, ans, st = alg.serve_trace(pairs.data_ptr(), True, ptrace)
    assert st["updates"] == PREFIX
    say(f"config4: device served {PREFIX} updates, {st['queries']} queries; oracle ingest")
    host = pairs[:PREFIX].cpu().numpy().astype(np.uint32)
    ref = oracle.CCAlg(N_VERT, SKETCH_SEED)
    ref.ingest_pairs(host, with_pre_insert=False)
    want = ref.dump()
    del ref
    got = alg.dump_sketches()
    sb = alg.sketch_bytes()
    assert got.size == want.size == N_VERT * sb
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} bytes differ, first in vertex {bad[0] // sb}"
    say("config4: prefix image equal to the oracle's")


def test_config4_every_answer_exact(workload):
    import torch

    pairs, trace = workload
    say("config4: edge lifetimes")
    life = exact_cc.EdgeLifetimes(pairs)
    alg = gz.CCSketchAlg(N_VERT, SKETCH_SEED)
    t0 = time.perf_counter()
    lat, ans, st = alg.serve_trace(pairs.data_ptr(), True, trace)
    say(f"config4: replay {time.perf_counter() - t0:.3f}s, {st['queries']} queries in {st['query_groups']} groups, "
        f"{st['boruvka_runs']} Borůvka runs")
    assert st["updates"] == N_UPD
    # query groups: the requests after each update batch, answered against the stream up to it
    upto, groups, cur = 0, [], None
    for i, r in enumerate(trace):
        if r["kind"] == gz.REQ_UPDATES:
            upto = int(r["offset"] + r["count"])
            cur = None
            continue
        if cur is None:
            cur = (upto, [])
            groups.append(cur)
        cur[1].append(i)
    assert len(groups) == st["query_groups"]
    t0 = time.time()
    for g, (p, idx) in enumerate(groups):
        ncc, lab = life.components(N_VERT, p)
        labh = lab.cpu().numpy()
        for i in idx:
            r = trace[i]
            want = ncc if r["kind"] == gz.REQ_CC else int(labh[r["u"]] == labh[r["v"]])
            assert ans[i] == want, f"group {g} at {p}: request {i} kind {r['kind']} got {ans[i]} want {want}"
        if g % 50 == 0:
            say(f"config4: group {g}/{len(groups)} at {p}: {ncc} components ({time.time() - t0:.0f}s)")
        del lab
    cc = alg.connected_components()
    ncc, lab = life.components(N_VERT, N_UPD)
    assert cc.size() == ncc
    assert np.array_equal(cc.canonical_labels(), lab.cpu().numpy().astype(np.uint32))
    say(f"config4: {len(groups)} groups exact, final partition exact ({ncc} components)")
    del life
    torch.cuda.empty_cache()
