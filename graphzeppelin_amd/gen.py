"""Seeded generators for the large stream configurations, run with torch on the device.

Streams of 10^8-10^9 updates with type-consistent deletes need a global dedup, which is a sort on
the GPU and far too slow on the host.  torch is only the array engine here (synthetic input
plumbing); nothing in the sketch path uses it.

Construction of a typed stream (no self loops, an edge is inserted at most once, deleted at most
once and only after its insert):
  1. draw candidate edges from the model (R-MAT or Chung-Lu) in a seeded order;
  2. drop self loops and keep the first occurrence of every edge, in draw order -> the inserts;
  3. pick n_del distinct inserts to delete, each scheduled at a uniform position between its
     insert and the end of the stream;
  4. order everything by schedule key.
Outputs are int32 (src, dst) pairs of shape (N, 2) and uint8 types (0 insert, 1 delete), on the
device the generator ran on, plus the final edge list when asked for.
"""
from __future__ import annotations

import math

import numpy as np
import torch

_CHUNK = 1 << 25


def _gen(seed: int, device) -> torch.Generator:
    g = torch.Generator(device=device)
    g.manual_seed(int(seed) & 0x7FFFFFFFFFFFFFFF)
    return g


def rmat_candidates(scale: int, count: int, a: float, b: float, c: float, g: torch.Generator, device,
                    n: int | None = None) -> torch.Tensor:
    """Edge keys (min << 32 | max, int64) of `count` R-MAT draws on 2^scale vertices; self loops, and
    draws with an endpoint >= n when n is given (R-MAT restricted to the first n vertices), are
    encoded as -1."""
    out = torch.empty(count, dtype=torch.int64, device=device)
    for s in range(0, count, _CHUNK):
        m = min(_CHUNK, count - s)
        src = torch.zeros(m, dtype=torch.int64, device=device)
        dst = torch.zeros(m, dtype=torch.int64, device=device)
        for _ in range(scale):
            r = torch.rand(m, generator=g, device=device)
            bit_s = (r >= a + b).to(torch.int64)
            bit_d = ((r >= a) & (r < a + b) | (r >= a + b + c)).to(torch.int64)
            src = src * 2 + bit_s
            dst = dst * 2 + bit_d
        lo, hi = torch.minimum(src, dst), torch.maximum(src, dst)
        key = (lo << 32) | hi
        bad = src == dst
        if n is not None:
            bad |= hi >= n
        key[bad] = -1
        out[s:s + m] = key
    return out


def chung_lu_weights(n: int, gamma: float, mean_deg: float) -> np.ndarray:
    """Expected degrees w_i = c (i + i0)^(-1/(gamma-1)) with mean `mean_deg` and the largest equal to
    sqrt(n * mean_deg) (the usual cap that keeps every pair probability <= 1)."""
    beta = 1.0 / (gamma - 1.0)
    idx = np.arange(n, dtype=np.float64)
    target = math.sqrt(n * mean_deg)

    def wmax(i0):
        s = np.sum((idx + i0) ** -beta)
        return mean_deg * n / s * i0 ** -beta

    lo, hi = 1e-3, float(n)
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        if wmax(mid) > target:
            lo = mid
        else:
            hi = mid
    i0 = math.sqrt(lo * hi)
    w = (idx + i0) ** -beta
    return w * (mean_deg * n / w.sum())


def chung_lu_candidates(n: int, count: int, gamma: float, mean_deg: float, g: torch.Generator, device) -> torch.Tensor:
    w = torch.from_numpy(chung_lu_weights(n, gamma, mean_deg)).to(device)
    cdf = torch.cumsum(w, 0)
    cdf = cdf / cdf[-1]
    perm = torch.randperm(n, generator=g, device=device)  # spread the hubs over the id space
    out = torch.empty(count, dtype=torch.int64, device=device)
    for s in range(0, count, _CHUNK):
        m = min(_CHUNK, count - s)
        u = torch.searchsorted(cdf, torch.rand(m, generator=g, device=device, dtype=torch.float64)).clamp_(max=n - 1)
        v = torch.searchsorted(cdf, torch.rand(m, generator=g, device=device, dtype=torch.float64)).clamp_(max=n - 1)
        u, v = perm[u], perm[v]
        lo, hi = torch.minimum(u, v), torch.maximum(u, v)
        key = (lo << 32) | hi
        key[u == v] = -1
        out[s:s + m] = key
    return out


def first_occurrences(keys: torch.Tensor) -> torch.Tensor:
    """Keys with self loops (-1) and repeats removed, first occurrence kept, draw order kept."""
    sk, si = torch.sort(keys, stable=True)
    first = torch.ones_like(sk, dtype=torch.bool)
    first[1:] = sk[1:] != sk[:-1]
    first &= sk >= 0
    keep = si[first]
    del sk, si, first
    keep, _ = torch.sort(keep)
    return keys[keep]


def typed_stream(ins_keys: torch.Tensor, n_del: int, g: torch.Generator, want_final: bool = False):
    """Inserts in order plus n_del deletes of distinct inserted edges, each scheduled uniformly
    between its insert and the end.  Returns (pairs int32 (N,2), types uint8 (N,), final keys or None)."""
    device = ins_keys.device
    n_ins = ins_keys.numel()
    n_del = min(n_del, n_ins)
    dj = torch.randperm(n_ins, generator=g, device=device)[:n_del]
    u = torch.rand(n_del, generator=g, device=device, dtype=torch.float64)
    dkey = dj.to(torch.float64) + 0.5 + u * (n_ins - dj.to(torch.float64))
    ikey = torch.arange(n_ins, device=device, dtype=torch.float64)
    order = torch.sort(torch.cat([ikey, dkey]), stable=True).indices
    del ikey, dkey, u
    which = torch.cat([torch.arange(n_ins, device=device), dj])[order]
    types = (order >= n_ins).to(torch.uint8)
    del order
    keys = ins_keys[which]
    swap = torch.rand(n_ins, generator=g, device=device) < 0.5  # orientation per edge
    sw = swap[which]
    lo = (keys >> 32).to(torch.int32)
    hi = (keys & 0xFFFFFFFF).to(torch.int32)
    pairs = torch.stack([torch.where(sw, hi, lo), torch.where(sw, lo, hi)], dim=1).contiguous()
    final = None
    if want_final:
        alive = torch.ones(n_ins, dtype=torch.bool, device=device)
        alive[dj] = False
        final = ins_keys[alive]
    return pairs, types, final


def _inserts(cand_fn, n_ins: int, g, device) -> torch.Tensor:
    draw = int(n_ins * 1.15) + 1024
    keys = first_occurrences(cand_fn(draw))
    while keys.numel() < n_ins:  # top up with more draws (keeps the first occurrence rule)
        more = cand_fn(max(1024, int((n_ins - keys.numel()) * 2)))
        keys = first_occurrences(torch.cat([keys, more]))
    return keys[:n_ins].contiguous()


def rmat_stream(scale: int, n_updates: int, del_frac: float, seed: int, device="cuda", a=0.57, b=0.19, c=0.19,
                want_final: bool = False, n: int | None = None):
    """R-MAT(a, b, c, d = 1-a-b-c) stream on 2^scale vertices (or its restriction to the first n
    of them, n <= 2^scale) with del_frac of the updates deleting present edges."""
    if n is not None and not (0 < n <= (1 << scale)):
        raise ValueError("n must lie in (0, 2^scale]")
    g = _gen(seed, device)
    n_del = int(round(n_updates * del_frac))
    n_ins = n_updates - n_del
    keys = _inserts(lambda cnt: rmat_candidates(scale, cnt, a, b, c, g, device, n), n_ins, g, device)
    return typed_stream(keys, n_del, g, want_final)


def chung_lu_stream(n: int, n_updates: int, del_frac: float, seed: int, device="cuda", gamma=2.1, mean_deg=64.0,
                    want_final: bool = False):
    """Chung-Lu stream: endpoints drawn independently with probability proportional to power-law
    expected degrees (exponent `gamma`, mean `mean_deg`, capped at sqrt(n * mean_deg))."""
    g = _gen(seed, device)
    n_del = int(round(n_updates * del_frac))
    n_ins = n_updates - n_del
    keys = _inserts(lambda cnt: chung_lu_candidates(n, cnt, gamma, mean_deg, g, device), n_ins, g, device)
    return typed_stream(keys, n_del, g, want_final)


def keys_to_pairs(keys: torch.Tensor) -> torch.Tensor:
    return torch.stack([(keys >> 32).to(torch.int32), (keys & 0xFFFFFFFF).to(torch.int32)], dim=1).contiguous()


def serving_trace(n: int, total: int, seed: int, per_request: int, bmin: int = 1000, bmax: int = 1_000_000,
                  cc_frac: float = 0.01, tenant: int = 0):
    """Request trace of the serving loop over a stream of `total` updates: update batches of
    log-uniform size in [bmin, bmax], each followed by Poisson(batch / per_request) query requests,
    a `cc_frac` share of them connected_components() and the rest u-v point queries on uniform
    vertex pairs.  Rows are gz_request records (graphzeppelin_amd.REQUEST_DTYPE), arrival times 0."""
    from ._lib import REQ_CC, REQ_POINT, REQ_UPDATES, REQUEST_DTYPE

    rng = np.random.default_rng(seed)
    rows = []
    off = 0
    while off < total:
        b = int(np.exp(rng.uniform(np.log(bmin), np.log(bmax))))
        b = min(b, total - off)
        rows.append((REQ_UPDATES, tenant, off, b, 0, 0, 0.0))
        off += b
        for _ in range(rng.poisson(b / per_request)):
            if rng.random() < cc_frac:
                rows.append((REQ_CC, tenant, 0, 0, 0, 0, 0.0))
            else:
                u, v = rng.integers(0, n, 2)
                rows.append((REQ_POINT, tenant, 0, 0, int(u), int(v), 0.0))
    return np.array(rows, dtype=REQUEST_DTYPE)


def trace_prefix(trace: np.ndarray, upto: int) -> np.ndarray:
    """The requests of `trace` that come before update `upto` (the last update batch clipped)."""
    out = []
    for r in trace:
        if r["kind"] == 0:
            if r["offset"] >= upto:
                break
            r = r.copy()
            r["count"] = min(int(r["count"]), upto - int(r["offset"]))
        out.append(r)
    return np.array(out, dtype=trace.dtype)


def multi_tenant_workload(num_tenants: int, seed: int, log2_lo: float = 12.0, log2_hi: float = 20.0,
                          updates_per_nlogn: float = 8.0, del_frac: float = 0.1, device="cuda"):
    """Tenants of the multi-tenant serving configuration: n log-uniform in [2^log2_lo, 2^log2_hi],
    each an R-MAT(0.57, 0.19, 0.19, 0.05) stream of updates_per_nlogn * n * log2(n) updates with
    del_frac deletes of present edges, drawn in chunks from torch's counter-based (Philox) generator
    seeded per tenant.  Returns [(n, pairs int32 (N, 2) on `device`)]."""
    rng = np.random.default_rng(seed)
    out = []
    for t in range(num_tenants):
        n = int(round(2.0 ** rng.uniform(log2_lo, log2_hi)))
        scale = max(1, math.ceil(math.log2(n)))
        m = int(round(updates_per_nlogn * n * math.log2(n)))
        pairs, _, _ = rmat_stream(scale, m, del_frac, seed * 1000 + t, device=device, n=n)
        out.append((n, pairs))
    return out


def multi_tenant_trace(tenants, seed: int, rate_updates_per_s: float | None, cc_per_tenant: float = 4.0,
                       point_frac: float = 0.5, bmin: int = 1000, bmax: int = 1_000_000):
    """Mixed request trace over tenants [(n, pairs)]: each tenant's stream arrives as update batches
    of log-uniform size in [bmin, bmax] in stream order; after a batch come Poisson-many query
    requests (on average cc_per_tenant connected_components() requests per tenant, at uniformly
    random stream positions, plus as many u-v point queries times point_frac / (1 - point_frac)),
    and every tenant ends with one connected_components() request.  Arrival times: every tenant's
    stream arrives at a steady rate over the same span, total updates / rate_updates_per_s, so the
    tenants together offer rate_updates_per_s throughout (None: everything arrives at time 0); a
    query arrives with the batch before it.  Requests are ordered by arrival time."""
    from ._lib import REQ_CC, REQ_POINT, REQ_UPDATES, REQUEST_DTYPE

    rng = np.random.default_rng(seed)
    rows = []
    span = None if rate_updates_per_s is None else sum(int(p.shape[0]) for _, p in tenants) / rate_updates_per_s
    for t, (n, pairs) in enumerate(tenants):
        total = int(pairs.shape[0])
        per_tenant_rate = None if span is None else total / span
        qrate = (cc_per_tenant / (1.0 - point_frac)) / total  # queries per update
        off = 0
        while off < total:
            b = min(int(np.exp(rng.uniform(np.log(bmin), np.log(bmax)))), total - off)
            arr = 0.0 if per_tenant_rate is None else (off + b) / per_tenant_rate
            rows.append((REQ_UPDATES, t, off, b, 0, 0, arr))
            off += b
            for _ in range(rng.poisson(b * qrate)):
                if rng.random() < point_frac:
                    u, v = rng.integers(0, n, 2)
                    rows.append((REQ_POINT, t, 0, 0, int(u), int(v), arr))
                else:
                    rows.append((REQ_CC, t, 0, 0, 0, 0, arr))
        rows.append((REQ_CC, t, 0, 0, 0, 0, 0.0 if per_tenant_rate is None else total / per_tenant_rate))
    arr = np.array(rows, dtype=REQUEST_DTYPE)
    order = np.argsort(arr["arrival_s"], kind="stable")  # stable: a tenant's requests keep their order
    return arr[order]

