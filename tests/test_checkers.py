"""The exact-components checker (tests/exact_cc.py) against the oracle's exact DSU on the CPU, before
the GPU tests trust it: typed R-MAT and Chung-Lu streams, components at many prefixes."""
import numpy as np
import pytest

import exact_cc


@pytest.mark.parametrize("model", ["rmat", "chung_lu"])
def test_lifetime_components_equal_exact_dsu(oracle_mod, model):
    import torch

    from graphzeppelin_amd import gen

    if model == "rmat":
        n = 1 << 10
        pairs, types, _ = gen.rmat_stream(10, 20_000, 0.3, 5, device="cpu")
    else:
        n = 3000
        pairs, types, _ = gen.chung_lu_stream(n, 8_000, 0.3, 6, device="cpu", mean_deg=3.0)
    host = pairs.numpy().astype(np.uint32)
    life = exact_cc.EdgeLifetimes(pairs)
    assert int((life.dele < life.N).sum()) == int(types.sum())  # every delete pairs with its insert
    for p in [0, 1, 17, 500, 4_000, pairs.shape[0] // 2, pairs.shape[0] - 1, pairs.shape[0]]:
        ncc, lab = life.components(n, p)
        c, want = oracle_mod.exact_cc(n, host[:p])
        assert ncc == c, p
        assert np.array_equal(lab.numpy().astype(np.uint32), want), p


def test_components_of_edges_on_a_path_and_isolated_vertices():
    import torch

    # a path 9-8-...-1 (labels must travel the whole path) plus isolated 0 and 10
    u = torch.arange(1, 9, dtype=torch.int64)
    v = u + 1
    ncc, lab = exact_cc.components_of_edges(11, v.flip(0), u.flip(0))
    assert ncc == 3
    assert lab.tolist() == [0] + [1] * 9 + [10]


def test_multi_tenant_trace_shape():
    from graphzeppelin_amd import gen
    import graphzeppelin_amd as gz

    tenants = gen.multi_tenant_workload(5, 3, log2_lo=8, log2_hi=10, device="cpu")
    for n, p in tenants:
        assert 256 <= n <= 1024 and int(p.max()) < n
        assert abs(p.shape[0] - 8 * n * np.log2(n)) <= 1
    trace = gen.multi_tenant_trace(tenants, 4, 1e6, bmin=10, bmax=1000)
    assert np.all(np.diff(trace["arrival_s"]) >= 0)
    for t, (n, p) in enumerate(tenants):
        sub = trace[trace["tenant"] == t]
        upd = sub[sub["kind"] == gz.REQ_UPDATES]
        # the stream arrives once, in order, at the tenant's steady rate
        assert upd["offset"][0] == 0 and np.array_equal(upd["offset"][1:], np.cumsum(upd["count"])[:-1])
        assert int(upd["count"].sum()) == p.shape[0]
        span = sum(q.shape[0] for _, q in tenants) / 1e6  # every stream arrives over the same span
        assert np.allclose(upd["arrival_s"], (upd["offset"] + upd["count"]) / (p.shape[0] / span))
        assert np.isclose(upd["arrival_s"][-1], span)
        assert sub[-1]["kind"] == gz.REQ_CC  # every stream ends with a connected_components() request
        pts = sub[sub["kind"] == gz.REQ_POINT]
        assert np.all(pts["u"] < n) and np.all(pts["v"] < n)
