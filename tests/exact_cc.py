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
