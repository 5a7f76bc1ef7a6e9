"""Host-fed ingestion on the GPU: every host path gives the oracle's sketches byte for byte, and the
file-driven harness (graphzeppelin_amd/gz_process_stream, the counterpart of the reference's
tools/process_stream.cpp) answers every query exactly with several reader threads.

Host paths: gz_alg_ingest_pairs from pageable memory (staged through pinned buffers), the serving
loop over a stream in pageable host memory and over one in pinned host memory (DMA'd in place), and
the device-resident serving loop, all on one Chung-Lu stream with churn."""
import json
import os
import subprocess

import numpy as np
import pytest

import exact_cc
import graphzeppelin_amd as gz
import oracle
from graphzeppelin_amd import streams

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(REPO, "graphzeppelin_amd", "gz_process_stream")
N_VERT = 1 << 14


def say(m):
    print(m, flush=True)


@pytest.fixture(scope="module")
def stream():
    from graphzeppelin_amd import gen

    pairs, types, _ = gen.chung_lu_stream(N_VERT, 3_000_000, 0.2, 9, device="cpu", mean_deg=32.0)
    return pairs.numpy().astype(np.uint32), types.numpy()


def test_host_ingest_paths_bit_exact(stream):
    import torch

    from graphzeppelin_amd import gen

    pairs, _ = stream
    N = pairs.shape[0]
    say(f"host ingest: {N} updates; oracle")
    ref = oracle.CCAlg(N_VERT, 5)
    ref.ingest_pairs(pairs, with_pre_insert=False)
    want = ref.dump()
    trace = gen.serving_trace(N_VERT, N, 3, 40_000, bmin=500, bmax=200_000)
    q = trace["kind"] != gz.REQ_UPDATES
    life = exact_cc.EdgeLifetimes(torch.from_numpy(pairs.astype(np.int32)))
    upto, exact = 0, np.zeros(trace.size, dtype=np.uint32)
    for i, r in enumerate(trace):
        if r["kind"] == gz.REQ_UPDATES:
            upto = int(r["offset"] + r["count"])
            continue
        c, lab = life.components(N_VERT, upto)
        exact[i] = c if r["kind"] == gz.REQ_CC else int(lab[r["u"]] == lab[r["v"]])
    # 1. the C ABI's host ingest, in pieces of uneven size (staging buffers are reused)
    a = gz.CCSketchAlg(N_VERT, 5)
    cuts = [0, 1, 777, 1_500_001, 1_500_002, N]
    for lo, hi in zip(cuts, cuts[1:]):
        a.ingest_pairs(pairs[lo:hi])
    assert np.array_equal(a.dump_sketches(), want)
    # 2. serving loop, pageable host stream
    b = gz.CCSketchAlg(N_VERT, 5)
    _, bans, _ = b.serve_trace(pairs.ctypes.data, False, trace)
    assert np.array_equal(bans[q], exact[q])
    assert np.array_equal(b.dump_sketches(), want)
    # 3. serving loop, pinned host stream (read in place by DMA)
    pinned = torch.from_numpy(pairs.astype(np.int32)).pin_memory()
    c = gz.CCSketchAlg(N_VERT, 5)
    _, cans, _ = c.serve_trace(pinned.data_ptr(), False, trace)
    assert np.array_equal(cans[q], exact[q])
    assert np.array_equal(c.dump_sketches(), want)
    # 4. serving loop, device-resident stream (grouped in place)
    dev = torch.from_numpy(pairs.astype(np.int32)).cuda()
    d = gz.CCSketchAlg(N_VERT, 5)
    _, dans, _ = d.serve_trace(dev.data_ptr(), True, trace)
    assert np.array_equal(dans[q], exact[q])
    assert np.array_equal(d.dump_sketches(), want)
    say(f"host ingest: 4 paths bit-exact, {int(q.sum())} answers exact")


@pytest.mark.parametrize("readers", [1, 4])
def test_process_stream_harness(stream, tmp_path, readers):
    import torch

    pairs, types = stream
    N = pairs.shape[0]
    path = str(tmp_path / "chung_lu.stream")
    streams.write_binary_stream(path, N_VERT, pairs, types)
    nq = 5
    say(f"harness: {N} updates, {nq} queries, {readers} readers")
    res = subprocess.run([HARNESS, path, str(nq), "1", str(readers), "42"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    out = json.loads(res.stdout.strip().splitlines()[-1])
    assert out["updates"] == N and out["updates_per_s"] > 0
    assert len(out["queries"]) == nq
    life = exact_cc.EdgeLifetimes(torch.from_numpy(pairs.astype(np.int32)))
    for q, rq in enumerate(out["queries"]):
        brk = int((q + 1) * (1.0 / nq) * N) if q + 1 < nq else N  # process_stream.cpp:115-119
        c, _ = life.components(N_VERT, brk)
        assert rq["cc"] == c, (q, brk)
        assert rq["total_s"] >= rq["flush_s"] >= 0 and rq["boruvka_s"] >= 0
    say(f"harness: {out['updates_per_s'] / 1e6:.1f} M updates/s with {readers} readers")
