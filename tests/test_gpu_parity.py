"""Parity of the HIP library with the CPU oracle, through the C ABI, on an MI355X.

Integer work, so the bar is bit-exact: every vertex's packed bucket array equals the oracle's byte
for byte, samples are equal, and component partitions equal the oracle's and an exact DSU over the
final edge set (compared as smallest-member labels, since DSU roots are arbitrary).
"""
import os

import numpy as np
import pytest

import graphzeppelin_amd as gz
import oracle
from graphzeppelin_amd import streams

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _directed_batches(pairs, chunk):
    """Both directions of every update grouped by source in stream order, cut into gutter-sized
    batches (the driver's two gutter inserts per update, include/graph_sketch_driver.h:171-172)."""
    by = {}
    for s, d in pairs.tolist():
        by.setdefault(s, []).append(d)
        by.setdefault(d, []).append(s)
    for v, ds in by.items():
        ds = np.asarray(ds, dtype=np.uint32)
        for i in range(0, ds.size, chunk):
            yield v, ds[i:i + chunk]


def _ingest(alg, pairs, path):
    if path == "host_pairs":
        alg.ingest_pairs(pairs)
    elif path == "device_pairs":
        import torch

        t = torch.from_numpy(np.ascontiguousarray(pairs.astype(np.int32))).cuda()
        torch.cuda.synchronize()
        alg.ingest_pairs_device(t.data_ptr(), pairs.shape[0])
        alg.prep_query()
        del t
    else:
        # the driver calls pre_insert on every update before gutter insertion
        # (include/graph_sketch_driver.h:169-172)
        for a, b in pairs.tolist():
            alg.pre_insert(a, b)
        for v, ds in _directed_batches(pairs, alg.get_desired_updates_per_batch()):
            alg.apply_update_batch(0, v, ds)


@pytest.mark.parametrize("gb", [4, 8])
@pytest.mark.parametrize("path", ["host_pairs", "device_pairs", "apply_batch", "plain"])
def test_config1_sketches_bit_exact(gb, path):
    say(f"config1 n=1024 gb={gb} path={path}")
    pairs, types = streams.erdos_renyi_churn_stream(1024, 0.1, 1)
    ref = oracle.CCAlg(1024, 42, gamma_bytes=gb)
    ref.ingest_pairs(pairs)
    cfg = gz.CCAlgConfiguration(gamma_bytes=gb, batched=(path != "plain"))
    alg = gz.CCSketchAlg(1024, 42, cfg)
    _ingest(alg, pairs, path)
    got = alg.dump_sketches()
    want = ref.dump()
    assert got.size == want.size
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"first mismatching byte {bad[:4]} (vertex {bad[0] // alg.sketch_bytes()})"
    cc = alg.connected_components()
    par, ncc = ref.connected_components()
    c, lab = oracle.exact_cc(1024, pairs)
    assert (cc.size(), ncc) == (c, c)
    assert np.array_equal(cc.canonical_labels(), lab)
    # force the Borůvka path on both sides and compare round counts
    nb_ref = oracle.CCAlg(1024, 42, gamma_bytes=gb, eager=False)
    nb_ref.ingest_pairs(pairs)
    nb_ref.connected_components()
    nb = gz.CCSketchAlg(1024, 42, gz.CCAlgConfiguration(gamma_bytes=gb, eager_dsu=False))
    nb.ingest_pairs(pairs)
    ccb = nb.connected_components()
    assert ccb.size() == c and np.array_equal(ccb.canonical_labels(), lab)
    assert nb.last_query_rounds == nb_ref.last_query_rounds


@pytest.mark.parametrize("gb", [4, 8])
def test_wide_sketches_bit_exact(gb):
    # more than 28 columns: the update kernel's write-back takes more than one pass over its LDS
    # delta (a separate kernel instance); images, partition and round count as the oracle's
    say(f"wide sketches gb={gb}")
    n, seed, sf = 300, 5, 3.0
    assert oracle.calc_cc_samples(n, sf) == 34
    pairs, _ = streams.erdos_renyi_churn_stream(n, 0.05, 11)
    ref = oracle.CCAlg(n, seed, sketches_factor=sf, gamma_bytes=gb, eager=False)
    ref.ingest_pairs(pairs)
    alg = gz.CCSketchAlg(n, seed, gz.CCAlgConfiguration(sketches_factor=sf, gamma_bytes=gb, eager_dsu=False))
    alg.ingest_pairs(pairs)
    assert np.array_equal(alg.dump_sketches(), ref.dump())
    cc = alg.connected_components()
    rpar, rc = ref.connected_components()
    c, lab = oracle.exact_cc(n, pairs)
    assert cc.size() == rc == c and np.array_equal(cc.canonical_labels(), lab)
    assert np.array_equal(cc.canonical_labels(), oracle.canonical(rpar))
    assert alg.last_query_rounds == ref.last_query_rounds
    # a standalone sketch of 35 columns, one vertex split over several tiles (atomic write-back)
    idx = np.random.default_rng(gb).integers(0, 1 << 30, size=3000, dtype=np.uint64)
    g = gz.Sketch(1 << 30, 77, 5, 7, gb)
    r = oracle.Sketch(1 << 30, 77, 5, 7, gb)
    g.update_batch(idx)
    r.update_batch(idx)
    assert g.serialize() == r.serialize()


@pytest.mark.parametrize("gb", [4, 8])
def test_standalone_sketch_parity(gb):
    say(f"standalone sketch gb={gb}")
    rng = np.random.default_rng(gb)
    for trial, (vl, samples, cols) in enumerate([(1 << 20, 6, 1), (4096, 3, 7), (100, 2, 7), (1 << 40, 25, 1)]):
        seed = int(rng.integers(0, 2**63))
        idx = rng.integers(0, vl, size=int(rng.integers(1, 5000)), dtype=np.uint64)
        g = gz.Sketch(vl, seed, samples, cols, gb)
        r = oracle.Sketch(vl, seed, samples, cols, gb)
        g.update_batch(idx)
        r.update_batch(idx)
        assert g.serialize() == r.serialize(), f"trial {trial}"
        for _ in range(samples):
            assert g.sample() == r.sample()
        with pytest.raises(gz.OutOfSamplesException):
            g.sample()


def test_sketch_known_answers():
    # test/sketch_test.cpp:14-18 and 66-80
    s = gz.Sketch(10, 77, 1, 7)
    assert s.sample()[1] == gz.ZERO
    s.update(1)
    with pytest.raises(gz.OutOfSamplesException):
        s.sample()
    s2 = gz.Sketch(40, 123, 1, 7)
    for _ in range(3):
        s2.update(0)
    assert s2.sample() == (0, gz.GOOD)


def test_sketch_merge_ops_parity():
    rng = np.random.default_rng(5)
    vl, seed, samples = 1 << 24, 99, 8
    a_idx = rng.integers(0, vl, 3000, dtype=np.uint64)
    b_idx = rng.integers(0, vl, 2000, dtype=np.uint64)
    ga, gb_ = gz.Sketch(vl, seed, samples), gz.Sketch(vl, seed, samples)
    ra, rb = oracle.Sketch(vl, seed, samples), oracle.Sketch(vl, seed, samples)
    ga.update_batch(a_idx); gb_.update_batch(b_idx); ra.update_batch(a_idx); rb.update_batch(b_idx)
    ga.range_merge(gb_, 3, 2)
    ra.range_merge(rb, 3, 2)
    assert ga.serialize() == ra.serialize()
    ga.merge(gb_)
    ra.merge(rb)
    assert ga.serialize() == ra.serialize()
    raw = rb.serialize()
    ga.merge_raw_bucket_buffer(raw)
    ra.merge_raw_bucket_buffer(raw)
    assert ga.serialize() == ra.serialize()
    # raw buffer merged twice cancels (test/sketch_test.cpp:423-458)
    z = gz.Sketch(4096, 7, 1, 1)
    s1 = gz.Sketch(4096, 7, 1, 1)
    s1.update_batch(np.arange(1024, dtype=np.uint64))
    r1 = s1.serialize()
    z.merge_raw_bucket_buffer(r1)
    i, res = z.sample()
    assert res != gz.ZERO
    z.merge_raw_bucket_buffer(r1)
    z.reset_sample_state()
    assert z.sample()[1] == gz.ZERO
    # serialize / deserialize round trip and equality (test/sketch_test.cpp:297-314)
    d = gz.Sketch(vl, seed, samples)
    d.deserialize(ga.serialize())
    assert d == ga
    ga.zero_contents()
    assert not any(ga.serialize())


def test_fixture_graph_78_components(golden_dir):
    # test/cc_alg_test.cpp:41-57
    n, edges = streams.read_ascii_graph(os.path.join(golden_dir, "multiples_graph_1024.txt"))
    alg = gz.CCSketchAlg(n, 2024)
    alg.ingest_pairs(edges)
    cc = alg.connected_components()
    assert cc.size() == 78
    _, lab = oracle.exact_cc(n, edges)
    assert np.array_equal(cc.canonical_labels(), lab)


def test_fixture_stream_queries_during_stream(golden_dir):
    # test/cc_alg_test.cpp:300-350: ten queries spaced through the binary stream
    n, pairs, types = streams.read_binary_stream(os.path.join(golden_dir, "multiples_graph_1024_stream.data"))
    alg = gz.CCSketchAlg(n, 31)
    step = pairs.shape[0] // 10
    for q in range(10):
        hi = pairs.shape[0] if q == 9 else (q + 1) * step
        alg.ingest_pairs(pairs[q * step:hi])
        alg.prep_query()
        cc = alg.connected_components()
        c, lab = oracle.exact_cc(n, pairs[:hi])
        assert cc.size() == c
        assert np.array_equal(cc.canonical_labels(), lab)
    assert cc.size() == 78


@pytest.mark.parametrize("seed", [1, 2])
def test_dynamic_streams_exact(seed):
    for density in (0.03, 0.002):
        pairs, _ = streams.dynamic_erdos_stream(1024, density, 0.5, 0.05, 3, seed)
        alg = gz.CCSketchAlg(1024, 500 + seed)
        alg.ingest_pairs(pairs)
        cc = alg.connected_components()
        c, lab = oracle.exact_cc(1024, pairs)
        assert cc.size() == c
        assert np.array_equal(cc.canonical_labels(), lab)


def test_point_query_grid():
    e = streams.grid_graph_edges(10, 10)
    e = e[~((e[:, 0] % 10 == 4) & (e[:, 1] % 10 == 5))]
    alg = gz.CCSketchAlg(100, 8)
    alg.ingest_pairs(e)
    cc = alg.connected_components()
    assert cc.size() == 2
    lab = cc.canonical_labels()
    us, vs = np.meshgrid(np.arange(10), np.arange(10))
    got = alg.point_queries(us.ravel(), vs.ravel())
    assert np.array_equal(got, lab[us.ravel()] == lab[vs.ravel()])
    for i in range(10):
        for j in range(10):
            assert alg.point_query(i, j) == (lab[i] == lab[j])


def test_eager_dsu_state_sequence():
    # test/cc_alg_test.cpp:223-263
    alg = gz.CCSketchAlg(100, 5)
    steps = [((1, 2), True), ((2, 3), True), ((1, 3), True), ((1, 3), True), ((2, 3), False), ((2, 3), True)]
    for (a, b), cached in steps:
        alg.update(a, b)
        assert alg.has_cached_query() == cached, (a, b)
        cc = alg.connected_components()
        assert alg.has_cached_query()
    lab = cc.canonical_labels()
    assert lab[1] == lab[2] == lab[3] and cc.size() == 98


def test_no_eager_dsu_always_runs_boruvka():
    alg = gz.CCSketchAlg(64, 5, gz.CCAlgConfiguration(eager_dsu=False))
    alg.update(1, 2)
    assert not alg.has_cached_query()
    assert alg.connected_components().size() == 63
    assert alg.has_cached_query()


@pytest.mark.parametrize("gb", [4, 8])
def test_write_binary_equals_oracle_checkpoint(tmp_path, gb):
    # write_binary / construct_from_serialized_data (src/cc_sketch_alg.cpp:614-623, 30-64) at a
    # non-default sketches_factor: the device's file equals the oracle's file byte for byte, and
    # both reloaded instances give the same round count and partition as the exact DSU
    n, seed, sf = 300, 1234, 2.5  # 28 columns instead of the default 18
    pairs, _ = streams.erdos_renyi_churn_stream(n, 0.05, 9)
    alg = gz.CCSketchAlg(n, seed, gz.CCAlgConfiguration(sketches_factor=sf, gamma_bytes=gb))
    alg.ingest_pairs(pairs)
    ref = oracle.CCAlg(n, seed, sketches_factor=sf, gamma_bytes=gb)
    ref.ingest_pairs(pairs)
    path, rpath = str(tmp_path / "alg.bin"), str(tmp_path / "ref.bin")
    alg.write_binary(path)
    ref.write_binary(rpath)
    raw, rraw = open(path, "rb").read(), open(rpath, "rb").read()
    assert int.from_bytes(raw[:8], "little") == seed and int.from_bytes(raw[8:12], "little") == n
    assert np.frombuffer(raw[12:20], dtype="<f8")[0] == sf
    assert alg.sketch_bytes() == (28 * 17 + 1) * (8 + gb) == ref.sketch_bytes()
    assert len(raw) == 20 + n * alg.sketch_bytes()
    assert raw == rraw
    back = gz.CCSketchAlg.construct_from_serialized_data(path, gz.CCAlgConfiguration(gamma_bytes=gb))
    rback = oracle.CCAlg.construct_from_serialized_data(rpath, gamma_bytes=gb)
    assert not back.has_cached_query() and not rback.has_cached_query()
    assert np.array_equal(back.dump_sketches(), rback.dump())
    cc = back.connected_components()
    rpar, rc = rback.connected_components()
    c, lab = oracle.exact_cc(n, pairs)
    assert cc.size() == rc == c
    assert np.array_equal(cc.canonical_labels(), oracle.canonical(rpar))
    assert np.array_equal(cc.canonical_labels(), lab)
    assert back.last_query_rounds == rback.last_query_rounds


def test_apply_raw_buckets_update_parity():
    pairs, _ = streams.erdos_renyi_churn_stream(128, 0.2, 4)
    ref = oracle.CCAlg(128, 6)
    ref.ingest_pairs(pairs)
    alg = gz.CCSketchAlg(128, 6)
    alg.ingest_pairs(pairs)
    raw = ref.dump(5, 1).tobytes()
    ref.apply_raw_buckets_update(9, raw)
    alg.apply_raw_buckets_update(9, raw)
    assert np.array_equal(alg.dump_sketches(), ref.dump())


def test_spanning_forest_is_a_forest_of_the_graph():
    pairs, _ = streams.erdos_renyi_churn_stream(400, 0.02, 12)
    alg = gz.CCSketchAlg(400, 3, gz.CCAlgConfiguration(eager_dsu=False))
    alg.ingest_pairs(pairs)
    sf = alg.calc_spanning_forest()
    c, lab = oracle.exact_cc(400, pairs)
    assert sf.shape[0] == 400 - c
    live = set(streams.edge_ids(streams.final_edge_set(pairs)).tolist())
    assert set(streams.edge_ids(sf).tolist()) <= live
    c2, lab2 = oracle.exact_cc(400, sf)
    assert c2 == c and np.array_equal(lab2, lab)


def test_majority_supernode_rounds_equal_oracle():
    # a giant component holding most vertices plus stragglers: rounds after the first few sum the
    # giant from its complement, which must give the oracle's partition and round count exactly
    pairs, _ = streams.dynamic_erdos_stream(1500, 0.0012, 0.3, 0.2, 3, 21)  # 59 components, largest 1431
    for gb in (4, 8):
        ref = oracle.CCAlg(1500, 77, gamma_bytes=gb, eager=False)
        ref.ingest_pairs(pairs)  # pre_insert without the eager DSU invalidates the cached answer
        alg = gz.CCSketchAlg(1500, 77, gz.CCAlgConfiguration(gamma_bytes=gb, eager_dsu=False))
        alg.ingest_pairs(pairs)
        cc = alg.connected_components()
        rparents, rc = ref.connected_components()
        assert alg.last_query_rounds == ref.last_query_rounds
        assert np.array_equal(cc.canonical_labels(), oracle.canonical(rparents))
        c, lab = oracle.exact_cc(1500, pairs)
        assert cc.size() == c and np.array_equal(cc.canonical_labels(), lab)


def test_one_sided_batches_and_raw_deltas_query_like_the_oracle():
    # apply_update_batch with one direction only and raw bucket deltas leave sketches whose XOR is
    # not zero; queries must then sum every supernode member-wise and still match the oracle
    pairs, _ = streams.erdos_renyi_churn_stream(700, 0.01, 5)
    ref = oracle.CCAlg(700, 31, eager=False)
    alg = gz.CCSketchAlg(700, 31, gz.CCAlgConfiguration(eager_dsu=False))
    ref.allocate_worker_memory(1)
    alg.allocate_worker_memory(1)
    for v, ds in _directed_batches(pairs, 64):
        keep = ds[(ds + v) % 7 != 0]  # drop some directions
        if keep.size:
            ref.apply_update_batch(0, v, keep)
            alg.apply_update_batch(0, v, keep)
    raw = ref.dump(3, 1).tobytes()
    ref.apply_raw_buckets_update(11, raw)
    alg.apply_raw_buckets_update(11, raw)
    ref.pre_insert(0, 1)  # without the eager DSU this invalidates the cached (empty) answer
    alg.pre_insert(0, 1)
    assert np.array_equal(alg.dump_sketches(), ref.dump())
    cc = alg.connected_components()
    rparents, rc = ref.connected_components()
    assert alg.last_query_rounds == ref.last_query_rounds
    assert np.array_equal(cc.canonical_labels(), oracle.canonical(rparents))


def test_query_graph_replays_equal_direct_launches():
    # the query prefix is replayed from a HIP graph; eager ingest and synchronisations run between
    # replays, and every replay must equal the same kernels launched one by one
    pairs, _ = streams.dynamic_erdos_stream(900, 0.003, 0.3, 0.2, 4, 8)
    g = gz.CCSketchAlg(900, 19, gz.CCAlgConfiguration(eager_dsu=False))
    d = gz.CCSketchAlg(900, 19, gz.CCAlgConfiguration(eager_dsu=False))
    d.set_query_graphs(False)
    step = pairs.shape[0] // 6 + 1
    for i in range(0, pairs.shape[0], step):
        chunk = pairs[i:i + step]
        g.ingest_pairs(chunk)
        d.ingest_pairs(chunk)
        cg, cd = g.connected_components(), d.connected_components()
        assert g.last_query_rounds == d.last_query_rounds
        assert np.array_equal(cg.canonical_labels(), cd.canonical_labels())
        c, lab = oracle.exact_cc(900, pairs[:i + step])
        assert cg.size() == c and np.array_equal(cg.canonical_labels(), lab)
        assert g.point_query(0, 1) == d.point_query(0, 1)


def test_round_finish_equals_launched_rounds():
    # Late rounds, once few vertices remain outside the giant, run inside one workgroup; queries
    # with that finish and with every round launched separately must agree round for round, and
    # with the exact components.  The first query sizes the next queries' launched rounds, so the
    # finish takes over from the second query on.
    pairs, _ = streams.dynamic_erdos_stream(1500, 0.0012, 0.3, 0.2, 3, 21)
    f = gz.CCSketchAlg(1500, 77, gz.CCAlgConfiguration(eager_dsu=False))
    l = gz.CCSketchAlg(1500, 77, gz.CCAlgConfiguration(eager_dsu=False))
    l.set_round_finish(False)
    took_over = 0
    step = pairs.shape[0] // 8 + 1
    for i in range(0, pairs.shape[0], step):
        chunk = pairs[i:i + step]
        f.ingest_pairs(chunk)
        l.ingest_pairs(chunk)
        cf, cl = f.connected_components(), l.connected_components()
        assert f.last_query_rounds == l.last_query_rounds
        assert f.last_query_round_sizes() == l.last_query_round_sizes()
        assert np.array_equal(cf.canonical_labels(), cl.canonical_labels())
        c, lab = oracle.exact_cc(1500, pairs[:i + step])
        assert cf.size() == c and np.array_equal(cf.canonical_labels(), lab)
        assert l.last_query_finish_round == 0
        if f.last_query_finish_round:
            took_over += 1
            assert 2 <= f.last_query_finish_round <= f.last_query_rounds
    assert took_over >= 2, took_over


def test_round_size_diagnostics():
    # per-round sizes of the last Borůvka query: one entry per round, at most n vertices regrouped,
    # and once a supernode is summed from its complement every later round walks only the vertices
    # regrouped the round before, so the counts never grow
    pairs, _ = streams.dynamic_erdos_stream(1500, 0.0012, 0.3, 0.2, 3, 21)
    alg = gz.CCSketchAlg(1500, 77, gz.CCAlgConfiguration(eager_dsu=False))
    alg.ingest_pairs(pairs)
    alg.connected_components()
    sizes = alg.last_query_round_sizes()
    assert len(sizes) == alg.last_query_rounds
    assert all(0 <= e <= 1500 for e, _ in sizes)
    elected = [i for i, (_, g) in enumerate(sizes) if g is not None]
    assert elected, sizes  # the largest component holds 1431 of 1500 vertices
    first = elected[0]
    assert all(g is not None for _, g in sizes[first:])
    assert sizes[first][0] <= 1500 - 2
    assert all(b[0] <= a[0] for a, b in zip(sizes[first:], sizes[first + 1:]))
