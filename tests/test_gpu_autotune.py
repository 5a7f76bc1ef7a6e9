"""The tile-size / flush-threshold autotuner (gz_autotune): every candidate is timed, the pick is the
fastest, and the chosen configuration changes how updates are cut into launches but not the
result -- sketches and components equal those of the default configuration."""
import numpy as np
import pytest

import graphzeppelin_amd as gz
from graphzeppelin_amd import gen

pytestmark = pytest.mark.gpu


def test_autotune_picks_fastest_and_keeps_results():
    scale, N = 14, 3_000_000
    n = 1 << scale
    pairs, _, _ = gen.rmat_stream(scale, N, 0.1, 11, device="cuda")
    cfg, trials = gz.autotune(n, 5, pairs.data_ptr(), N)
    assert len(trials) == 12  # 4 tile sizes x 3 flush thresholds
    assert all(ms > 0 for _, _, ms in trials)
    best = min(trials, key=lambda t: t[2])
    assert (cfg.tile, cfg.flush_threshold) == best[:2]
    assert cfg.eager_dsu and cfg.batched  # the caller's (default) choices are kept
    tuned = gz.CCSketchAlg(n, 5, cfg)
    plain = gz.CCSketchAlg(n, 5)
    for lo in range(0, N, 1_000_000):
        tuned.ingest_pairs_device(pairs[lo:lo + 1_000_000].data_ptr(), min(1_000_000, N - lo))
        plain.ingest_pairs_device(pairs[lo:lo + 1_000_000].data_ptr(), min(1_000_000, N - lo))
        a, b = tuned.connected_components(), plain.connected_components()
        assert np.array_equal(a.canonical_labels(), b.canonical_labels())
    assert tuned.digest() == plain.digest()
    assert np.array_equal(tuned.dump_sketches(0, 256), plain.dump_sketches(0, 256))
