"""Graph-stream formats and seeded stream generators (the StreamingUtilities role).

* The binary stream format (README.md:57-72 of the reference): a 4-byte vertex count, an 8-byte
  update count, then 9-byte records {uint8 type, uint32 src, uint32 dst}, type 0 = insert,
  1 = delete.
* The ASCII graph format of test/res/multiples_graph_1024.txt: "n m" then "src dst" lines.
* Generators emit type-consistent streams (no self loops, no insert of a present edge, deletes only
  of present edges) as (src, dst) uint32 pairs plus a uint8 type array.  The generators here are
  numpy (PCG64 seeded); the 10^8-update ones run on the device with torch (graphzeppelin_amd/gen.py).
"""
from __future__ import annotations

import numpy as np

REC = np.dtype([("type", "u1"), ("src", "<u4"), ("dst", "<u4")])


def write_binary_stream(path: str, num_vertices: int, pairs: np.ndarray, types: np.ndarray | None = None) -> None:
    pairs = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
    rec = np.empty(pairs.shape[0], dtype=REC)
    rec["type"] = 0 if types is None else np.asarray(types, dtype=np.uint8)
    rec["src"] = pairs[:, 0]
    rec["dst"] = pairs[:, 1]
    with open(path, "wb") as f:
        f.write(np.uint32(num_vertices).tobytes())
        f.write(np.uint64(pairs.shape[0]).tobytes())
        f.write(rec.tobytes())


def read_binary_stream(path: str) -> tuple[int, np.ndarray, np.ndarray]:
    """Returns (num_vertices, pairs (m,2) uint32, types (m,) uint8)."""
    with open(path, "rb") as f:
        hdr = f.read(12)
        n = int(np.frombuffer(hdr[:4], dtype="<u4")[0])
        m = int(np.frombuffer(hdr[4:12], dtype="<u8")[0])
        rec = np.frombuffer(f.read(m * REC.itemsize), dtype=REC)
    if rec.shape[0] != m:
        raise ValueError(f"{path}: truncated stream ({rec.shape[0]} of {m} records)")
    pairs = np.stack([rec["src"], rec["dst"]], axis=1).astype(np.uint32)
    return n, pairs, rec["type"].copy()


def read_ascii_graph(path: str) -> tuple[int, np.ndarray]:
    """Returns (num_vertices, edges (m,2) uint32) of an ASCII edge-list file."""
    with open(path) as f:
        n, m = (int(x) for x in f.readline().split())
        data = np.loadtxt(f, dtype=np.int64, ndmin=2)
    if data.shape[0] != m:
        raise ValueError(f"{path}: expected {m} edges, found {data.shape[0]}")
    return n, data[:, :2].astype(np.uint32)


def edge_ids(pairs: np.ndarray) -> np.ndarray:
    """concat_pairing_fn (src/util.cpp:45-51): (min << 32) | max."""
    p = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    lo = np.minimum(p[:, 0], p[:, 1])
    hi = np.maximum(p[:, 0], p[:, 1])
    return (lo << np.uint64(32)) | hi


def final_edge_set(pairs: np.ndarray) -> np.ndarray:
    """Edges whose update count is odd (the toggle semantics of the sketches)."""
    ids = edge_ids(pairs)
    u, c = np.unique(ids, return_counts=True)
    e = u[(c & 1) == 1]
    return np.stack([(e >> np.uint64(32)).astype(np.uint32), (e & np.uint64(0xFFFFFFFF)).astype(np.uint32)], axis=1)


def check_type_consistent(pairs: np.ndarray, types: np.ndarray) -> None:
    """Raise if the stream inserts a present edge, deletes an absent one or has a self loop."""
    pairs = np.asarray(pairs).reshape(-1, 2)
    if np.any(pairs[:, 0] == pairs[:, 1]):
        raise ValueError("self loop in stream")
    ids = edge_ids(pairs)
    order = np.argsort(ids, kind="stable")
    sid, st = ids[order], np.asarray(types)[order]
    first = np.ones(sid.size, dtype=bool)
    first[1:] = sid[1:] != sid[:-1]
    # within each edge group the types must alternate starting with insert
    grp_start = np.maximum.accumulate(np.where(first, np.arange(sid.size), 0))
    pos = np.arange(sid.size) - grp_start
    if np.any(st != (pos & 1)):
        raise ValueError("stream is not type-consistent")


def erdos_renyi_churn_stream(n: int = 1024, p: float = 0.1, seed: int = 1, delete_frac: float = 0.2,
                             reinsert_frac: float = 0.1) -> tuple[np.ndarray, np.ndarray]:
    """G(n, p) over all vertex pairs; every edge inserted in seeded random order, then
    `delete_frac` of the edges deleted and `reinsert_frac` of those reinserted, deletes and
    reinserts interleaved at random with each reinsert after its delete."""
    rng = np.random.Generator(np.random.PCG64(seed))
    iu, ju = np.triu_indices(n, k=1)
    keep = rng.random(iu.size) < p
    edges = np.stack([iu[keep], ju[keep]], axis=1).astype(np.uint32)
    edges = edges[rng.permutation(edges.shape[0])]
    swap = rng.random(edges.shape[0]) < 0.5
    edges[swap] = edges[swap][:, ::-1]
    m = edges.shape[0]
    n_del = int(round(delete_frac * m))
    dele = rng.choice(m, size=n_del, replace=False)
    n_re = int(round(reinsert_frac * n_del))
    re = dele[rng.choice(n_del, size=n_re, replace=False)]
    # interleave: deletes get random keys in [0,1); a reinsert gets a key after its delete
    dkey = rng.random(n_del)
    pos_of = dict(zip(dele.tolist(), dkey.tolist()))
    rkey = np.array([pos_of[int(e)] + (1 - pos_of[int(e)]) * rng.random() for e in re])
    keys = np.concatenate([dkey, rkey])
    tail_edges = np.concatenate([edges[dele], edges[re]])
    tail_types = np.concatenate([np.ones(n_del, np.uint8), np.zeros(n_re, np.uint8)])
    order = np.argsort(keys, kind="stable")
    pairs = np.concatenate([edges, tail_edges[order]])
    types = np.concatenate([np.zeros(m, np.uint8), tail_types[order]])
    return pairs, types


def dynamic_erdos_stream(n: int, density: float, delete_portion: float, adtl_portion: float, rounds: int,
                         seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Multi-round dynamic Erdős–Rényi stream in the spirit of the reference tests'
    DynamicErdosGenerator(seed, n, density, delete_portion, adtl_portion, rounds)
    (test/cc_alg_test.cpp:21-33): each round inserts `density` of the absent pairs, then deletes
    `delete_portion` of the present edges and re-adds `adtl_portion` of those."""
    rng = np.random.Generator(np.random.PCG64(seed))
    iu, ju = np.triu_indices(n, k=1)
    present = np.zeros(iu.size, dtype=bool)
    out_p, out_t = [], []
    for _ in range(rounds):
        absent = np.flatnonzero(~present)
        ins = absent[rng.random(absent.size) < density]
        ins = ins[rng.permutation(ins.size)]
        present[ins] = True
        out_p.append(np.stack([iu[ins], ju[ins]], 1))
        out_t.append(np.zeros(ins.size, np.uint8))
        pres = np.flatnonzero(present)
        dele = pres[rng.random(pres.size) < delete_portion]
        dele = dele[rng.permutation(dele.size)]
        present[dele] = False
        out_p.append(np.stack([iu[dele], ju[dele]], 1))
        out_t.append(np.ones(dele.size, np.uint8))
        re = dele[rng.random(dele.size) < adtl_portion]
        present[re] = True
        out_p.append(np.stack([iu[re], ju[re]], 1))
        out_t.append(np.zeros(re.size, np.uint8))
    return np.concatenate(out_p).astype(np.uint32), np.concatenate(out_t)


def grid_graph_edges(rows: int, cols: int) -> np.ndarray:
    v = np.arange(rows * cols).reshape(rows, cols)
    h = np.stack([v[:, :-1].ravel(), v[:, 1:].ravel()], 1)
    w = np.stack([v[:-1, :].ravel(), v[1:, :].ravel()], 1)
    return np.concatenate([h, w]).astype(np.uint32)


def _tri_pairs(s: int, k: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Linear index k in [0, s(s-1)/2) -> (i, j), i < j, row-major over the upper triangle."""
    k = k.astype(np.float64)
    i = (s - 2 - np.floor(np.sqrt(-8 * k + 4.0 * s * (s - 1) - 7) / 2 - 0.5)).astype(np.int64)
    ki = k.astype(np.int64)
    j = ki + i + 1 - s * (s - 1) // 2 + (s - i) * ((s - i) - 1) // 2
    return i, j


def planted_blocks_stream(n: int = 1 << 15, nblocks: int = 64, p: float = 0.05, churn: float = 0.3, seed: int = 2,
                          small: tuple[int, int] = (16, 128)):
    """Planted components: one giant block plus nblocks-1 blocks of uniform size in `small`, vertex
    ids assigned by a seeded permutation; G(size, p) edges inside every block (none across); every
    edge inserted in seeded random order, then `churn` of the edges deleted and reinserted, each
    reinsert after its delete.  Returns (pairs, types, block_of_vertex)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.integers(small[0], small[1] + 1, nblocks - 1)
    sizes = np.concatenate([[n - int(sizes.sum())], sizes]).astype(np.int64)
    perm = rng.permutation(n).astype(np.uint32)
    block = np.empty(n, dtype=np.int32)
    edges = []
    start = 0
    for b, s in enumerate(sizes):
        ids = perm[start:start + s]
        block[ids] = b
        start += s
        m = s * (s - 1) // 2
        if m == 0:
            continue
        pos = []
        cur = -1
        while cur < m:
            gaps = rng.geometric(p, size=int(max(64, (m - cur) * p * 1.1)))
            run = cur + np.cumsum(gaps)
            pos.append(run[run < m])
            cur = int(run[-1])
        k = np.concatenate(pos)
        i, j = _tri_pairs(int(s), k)
        edges.append(np.stack([ids[i], ids[j]], axis=1))
    e = np.concatenate(edges).astype(np.uint32)
    e = e[rng.permutation(e.shape[0])]
    swap = rng.random(e.shape[0]) < 0.5
    e[swap] = e[swap][:, ::-1]
    m = e.shape[0]
    nc = int(round(churn * m))
    ch = rng.choice(m, size=nc, replace=False)
    dkey = rng.random(nc)
    rkey = dkey + (1 - dkey) * rng.random(nc)
    keys = np.concatenate([dkey, rkey])
    tail = np.concatenate([e[ch], e[ch]])
    ttypes = np.concatenate([np.ones(nc, np.uint8), np.zeros(nc, np.uint8)])
    order = np.argsort(keys, kind="stable")
    pairs = np.concatenate([e, tail[order]])
    types = np.concatenate([np.zeros(m, np.uint8), ttypes[order]])
    return pairs, types, block
