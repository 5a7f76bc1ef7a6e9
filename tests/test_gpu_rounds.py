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


def _check_query(alg, ref, n, seed):
    vl = oracle.calc_vector_length(n)
    D = oracle.calc_bkt_per_col(vl)
    C = oracle.calc_cc_samples(n, 1.0)
    alpha, gamma = parse_images(ref.dump(), n, C * D + 1)
    parts = partitions(alpha, gamma, D, seed, C)
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
