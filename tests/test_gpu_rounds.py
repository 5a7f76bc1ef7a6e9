"""Borůvka round by round: the device's per-round diagnostics against a CPU restatement.

The partition after every round of boruvka_emulation (src/cc_sketch_alg.cpp:464-513) is fixed by
the sketches.  tests/boruvka_rounds.py recomputes those partitions from the oracle's sketch images;
the device reports, per round, how many vertices it regrouped and the root of the giant supernode
it summed from its complement.  Without a giant every vertex is regrouped; with one, exactly the
vertices outside the giant's component after the previous round are.  The round count and the final
partition must match as well.  Both the launched rounds and the single-workgroup finish of the last
rounds are checked.
"""
import numpy as np
import pytest

import graphzeppelin_amd as gz
import oracle
from graphzeppelin_amd import gen, streams

from boruvka_rounds import parse_images, partitions

pytestmark = pytest.mark.gpu


def _check_query(alg, ref, n, seed, gb=4):
    vl = oracle.calc_vector_length(n)
    D = oracle.calc_bkt_per_col(vl)
    C = oracle.calc_cc_samples(n, 1.0)
    alpha, gamma = parse_images(ref.dump(), n, C * D + 1, gb)
    parts = partitions(alpha, gamma, D, seed, C, gb)
    cc = alg.connected_components()
    rounds = alg.last_query_rounds
    assert rounds == len(parts) - 1
    assert np.array_equal(cc.canonical_labels(), oracle.canonical(parts[-1]))
    sizes = alg.last_query_round_sizes()
    assert len(sizes) == rounds
    for r, (regrouped, giant) in enumerate(sizes, start=1):
        if giant is None:
            assert regrouped == n, (r, sizes)
        else:
            prev = parts[r - 1]
            assert regrouped == n - int(np.count_nonzero(prev == prev[giant])), (r, sizes)
    return alg.last_query_finish_round


@pytest.mark.parametrize("finish", [True, False])
def test_rounds_match_cpu_partitions_erdos(finish):
    n, seed = 1500, 77
    pairs, _ = streams.dynamic_erdos_stream(n, 0.0012, 0.3, 0.2, 3, 21)
    alg = gz.CCSketchAlg(n, seed, gz.CCAlgConfiguration(eager_dsu=False))
    alg.set_round_finish(finish)
    ref = oracle.CCAlg(n, seed, eager=False)
    took_over = 0
    step = pairs.shape[0] // 8 + 1
    for i in range(0, pairs.shape[0], step):
        alg.ingest_pairs(pairs[i:i + step])
        ref.ingest_pairs(pairs[i:i + step])
        if _check_query(alg, ref, n, seed):
            took_over += 1
    assert (took_over > 0) == finish


def _complete_bipartite(m):
    left = np.repeat(np.arange(m, dtype=np.uint32), m)
    right = np.tile(np.arange(m, 2 * m, dtype=np.uint32), m)
    return np.stack([left, right], axis=1)


@pytest.mark.parametrize("gb,deep_skip", [(4, True), (4, False), (8, True)])
def test_rounds_match_cpu_partitions_deep_buckets(gb, deep_skip, tmp_path):
    # K_{1024,1024}: after round 1 the supernodes hold tens to hundreds of vertices of degree 1024,
    # so their cuts run to 10^4..10^5 edges and their samples come from depths of 14 and more:
    # records past the first line of a column's overflow run (depth >= 15 with 12-byte records),
    # which a reduce reads only for members whose deep flag is set (about 3% of them here, at
    # 1024 updates per vertex).  Every round runs as its own reduce (finish off) and must give the
    # CPU restatement's partition; so must a reloaded checkpoint (flags set by the overwrite) and
    # an instance built from raw bucket images (flags set by the XOR merge).  The sketch images
    # stay bit-exact: the flags live outside the records.  Rows of 16-byte records keep no flags
    # (byte 60 of a half-line is inside a record there) and read every record.
    n, m, seed = 2048, 1024, 9
    pairs = _complete_bipartite(m)
    ref = oracle.CCAlg(n, seed, gamma_bytes=gb, eager=False)
    ref.ingest_pairs(pairs)
    vl = oracle.calc_vector_length(n)
    D = oracle.calc_bkt_per_col(vl)
    C = oracle.calc_cc_samples(n, 1.0)
    alpha, gamma = parse_images(ref.dump(), n, C * D + 1, gb)
    depths = []
    partitions(alpha, gamma, D, seed, C, gb, depths=depths)
    assert max(max(d, default=0) for d in depths[1:]) >= 15, depths
    cfg = gz.CCAlgConfiguration(eager_dsu=False, gamma_bytes=gb)

    def check(alg):
        alg.set_round_finish(False)
        alg.set_deep_skip(deep_skip)
        assert np.array_equal(alg.dump_sketches(), ref.dump())
        assert _check_query(alg, ref, n, seed, gb) == 0

    alg = gz.CCSketchAlg(n, seed, cfg)
    alg.ingest_pairs(pairs)
    check(alg)
    path = str(tmp_path / "alg.bin")
    alg.write_binary(path)
    check(gz.CCSketchAlg.construct_from_serialized_data(path, cfg))
    raw = gz.CCSketchAlg(n, seed, cfg)
    images = ref.dump().reshape(n, -1)
    for v in range(n):
        raw.apply_raw_buckets_update(v, images[v].tobytes())
    raw.pre_insert(0, m)  # without the eager DSU this invalidates the cached (empty) answer
    check(raw)


def test_rounds_match_cpu_partitions_power_law():
    # a heavy-tailed graph: hubs make large supernodes early, so candidate giants (a supernode with
    # at least n/32 vertices, excluded a round before it holds a majority) are exercised
    n, seed = 4096, 5
    p, _, _ = gen.chung_lu_stream(n, 120_000, 0.15, 3, device="cpu", mean_deg=6.0)
    pairs = p.numpy().astype(np.uint32).reshape(-1, 2)
    alg = gz.CCSketchAlg(n, seed, gz.CCAlgConfiguration(eager_dsu=False))
    ref = oracle.CCAlg(n, seed, eager=False)
    step = pairs.shape[0] // 4 + 1
    for i in range(0, pairs.shape[0], step):
        alg.ingest_pairs(pairs[i:i + step])
        ref.ingest_pairs(pairs[i:i + step])
        _check_query(alg, ref, n, seed)
