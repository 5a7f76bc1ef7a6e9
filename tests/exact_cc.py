"""Exact connected components used as the checker of GPU answers (test infrastructure).

Independent of the library's kernels: scipy's connected_components on the host, or min-label
propagation with pointer jumping in torch on the device for streams too large for the host.
"""
from __future__ import annotations

import numpy as np
import torch


def exact_components(n: int, edges: np.ndarray) -> tuple[int, np.ndarray]:
    """Connected components of an edge list with scipy (independent of this library); labels are
    the smallest vertex id of each component."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
    m = sp.coo_matrix((np.ones(e.shape[0], dtype=np.int8), (e[:, 0], e[:, 1])), shape=(n, n)).tocsr()
    ncc, lab = connected_components(m, directed=False)
    first = np.full(ncc, n, dtype=np.int64)
    np.minimum.at(first, lab, np.arange(n))
    return int(ncc), first[lab].astype(np.uint32)


def exact_components_torch(n: int, keys: torch.Tensor) -> tuple[int, torch.Tensor]:
    """Components of the edges `keys` (min << 32 | max) by min-label propagation with pointer
    jumping, in torch on the keys' device (independent of this library's kernels).  Labels are the
    smallest vertex id of each component."""
    device = keys.device
    u = (keys >> 32).to(torch.int64)
    v = (keys & 0xFFFFFFFF).to(torch.int64)
    lab = torch.arange(n, device=device, dtype=torch.int64)
    while True:
        prev = lab.clone()
        m = torch.minimum(lab[u], lab[v])
        lab.scatter_reduce_(0, u, m, reduce="amin")
        lab.scatter_reduce_(0, v, m, reduce="amin")
        while True:  # pointer jumping
            nxt = lab[lab]
            if torch.equal(nxt, lab):
                break
            lab = nxt
        if torch.equal(lab, prev):
            break
    ncc = int((lab == torch.arange(n, device=device)).sum().item())
    return ncc, lab


def prefix_final_keys(pairs: torch.Tensor, upto: int) -> torch.Tensor:
    """Edges present after the first `upto` updates of a typed stream (odd update count)."""
    p = pairs[:upto].to(torch.int64)
    lo = torch.minimum(p[:, 0], p[:, 1])
    hi = torch.maximum(p[:, 0], p[:, 1])
    k, c = torch.unique((lo << 32) | hi, return_counts=True)
    return k[(c & 1) == 1]


def components_of_edges(n: int, u: torch.Tensor, v: torch.Tensor) -> tuple[int, torch.Tensor]:
    """Components of the edges (u[i], v[i]) (int64 index tensors on one device) by min-label
    propagation with pointer jumping, int32 labels.  Labels are the smallest vertex id of each
    component: a label is always a vertex of the same component, it only decreases, and it stops
    changing only once every edge has equal labels at both ends."""
    device = u.device
    lab = torch.arange(n, device=device, dtype=torch.int32)
    while True:
        prev = lab.clone()
        m = torch.minimum(lab[u], lab[v])
        lab.scatter_reduce_(0, u, m, reduce="amin")
        lab.scatter_reduce_(0, v, m, reduce="amin")
        del m
        while True:  # pointer jumping
            nxt = lab[lab.long()]
            if torch.equal(nxt, lab):
                break
            lab = nxt
        if torch.equal(lab, prev):
            break
    ncc = int((lab == torch.arange(n, device=device, dtype=torch.int32)).sum().item())
    return ncc, lab


class EdgeLifetimes:
    """Live edge sets of a typed stream (every edge inserted at most once and deleted at most once,
    after its insert) at any prefix, on the device: one stable sort of the N edge keys gives each
    edge its insert position and its delete position (N when it is never deleted); the edges live
    after the first p updates are those inserted before p and not deleted before p."""

    def __init__(self, pairs: torch.Tensor):
        p = pairs.to(torch.int64)
        lo = torch.minimum(p[:, 0], p[:, 1])
        hi = torch.maximum(p[:, 0], p[:, 1])
        del p
        keys = (lo << 32) | hi
        del lo, hi
        self.N = keys.numel()
        sk, order = torch.sort(keys, stable=True)
        del keys
        start = torch.ones_like(sk, dtype=torch.bool)
        start[1:] = sk[1:] != sk[:-1]
        first = torch.nonzero(start).squeeze(1)
        del start
        count = torch.diff(first, append=torch.tensor([self.N], device=first.device))
        if int(count.max()) > 2:
            raise ValueError("an edge occurs more than twice: the stream is not type-consistent")
        self.ins = order[first]
        second = torch.clamp(first + 1, max=self.N - 1)
        self.dele = torch.where(count == 2, order[second], torch.full_like(first, self.N))
        k = sk[first]
        del sk, order, second, count, first
        self.u = k >> 32
        self.v = k & 0xFFFFFFFF

    def live(self, p: int) -> tuple[torch.Tensor, torch.Tensor]:
        m = (self.ins < p) & (self.dele >= p)
        return self.u[m], self.v[m]

    def components(self, n: int, p: int) -> tuple[int, torch.Tensor]:
        u, v = self.live(p)
        return components_of_edges(n, u, v)
