"""Parity at the stream configurations' full sizes (GPU).

* planted blocks, n = 2^15, ~31M updates with churn, 64-bit checksums: every vertex sketch equals
  the oracle's (digest over all bytes plus full-row byte checks on sampled vertices) and the
  partition equals an exact DSU;
* R-MAT(0.57, 0.19, 0.19, 0.05), n = 2^17, 100M updates with 10% deletes: a connected-components
  query every 10M updates, each partition equal to exact components of that prefix;
* the 9-byte binary stream format read back and ingested.
"""
import os

import numpy as np
import pytest

import exact_cc
import graphzeppelin_amd as gz
import oracle
from graphzeppelin_amd import streams

pytestmark = pytest.mark.gpu


def say(m):
    print(m, flush=True)


def test_planted_blocks_2e15_bit_exact():
    n = 1 << 15
    say("planted blocks: generating")
    pairs, types, block = streams.planted_blocks_stream(n, 64, 0.05, 0.3, 2)
    say(f"planted blocks: {pairs.shape[0]} updates; oracle ingest")
    ref = oracle.CCAlg(n, 42, gamma_bytes=8)
    ref.ingest_pairs(pairs, with_pre_insert=False)
    say("planted blocks: device ingest")
    alg = gz.CCSketchAlg(n, 42, gz.CCAlgConfiguration(gamma_bytes=8))
    alg.ingest_pairs(pairs)
    img = ref.dump()
    nb = alg.sketch_bytes() // 16
    assert alg.digest() == gz.digest_of_image(img, nb, 8)
    sb = alg.sketch_bytes()
    rng = np.random.default_rng(0)
    for v in rng.choice(n, 64, replace=False).tolist() + [0, n - 1]:
        assert np.array_equal(alg.dump_sketches(v, 1), img[v * sb:(v + 1) * sb]), v
    say("planted blocks: components")
    cc = alg.connected_components()
    c, lab = oracle.exact_cc(n, pairs)
    assert cc.size() == c
    assert np.array_equal(cc.canonical_labels(), lab)
    # no edge crosses blocks, so every component lies inside one block
    assert np.all(block[cc.canonical_labels()] == block)


def test_rmat_2e17_queries_every_10m():
    import torch

    from graphzeppelin_amd import gen

    scale, N = 17, 100_000_000
    n = 1 << scale
    say("rmat: generating")
    pairs, types, _ = gen.rmat_stream(scale, N, 0.1, 3, device="cuda")
    # tile size and flush threshold autotuned on the first 10M updates of the stream
    cfg, trials = gz.autotune(n, 77, pairs.data_ptr(), 10_000_000)
    say(f"rmat: autotuned tile {cfg.tile} flush threshold {cfg.flush_threshold} from {len(trials)} trials")
    assert (cfg.tile, cfg.flush_threshold) in [(t, f) for t, f, _ in trials]
    alg = gz.CCSketchAlg(n, 77, cfg)
    step = 10_000_000
    for q in range(10):
        lo, hi = q * step, (q + 1) * step
        alg.ingest_pairs_device(pairs[lo:hi].data_ptr(), hi - lo)
        cc = alg.connected_components()
        keys = exact_cc.prefix_final_keys(pairs, hi)
        c, lab = exact_cc.exact_components_torch(n, keys)
        say(f"rmat: query {q} at {hi}: components {cc.size()} exact {c} rounds {alg.last_query_rounds}")
        assert cc.size() == c
        assert np.array_equal(cc.canonical_labels(), lab.cpu().numpy().astype(np.uint32))
        del keys, lab
        torch.cuda.empty_cache()


def test_binary_stream_file_round_trip(tmp_path):
    pairs, types = streams.erdos_renyi_churn_stream(1024, 0.1, 1)
    path = str(tmp_path / "c1.stream")
    streams.write_binary_stream(path, 1024, pairs, types)
    assert os.path.getsize(path) == 12 + 9 * pairs.shape[0]
    n, p2, t2 = streams.read_binary_stream(path)
    assert n == 1024 and np.array_equal(p2, pairs) and np.array_equal(t2, types)
    alg = gz.CCSketchAlg(n, 42)
    alg.ingest_pairs(p2)
    ref = oracle.CCAlg(n, 42)
    ref.ingest_pairs(pairs)
    assert np.array_equal(alg.dump_sketches(), ref.dump())
