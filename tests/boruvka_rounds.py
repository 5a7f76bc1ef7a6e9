"""Round-by-round Borůvka on the CPU over the oracle's sketch images (test infrastructure).

boruvka_emulation (src/cc_sketch_alg.cpp:464-513) is deterministic at the level of partitions:
round 0 samples column 0 of every vertex sketch (run_round_zero, 223-244), round r samples column r
of every supernode's merged sketch (perform_boruvka_round, 246-380; range_merge + sample,
src/sketch.cpp:94-179), and the partition after a round is the closure of the previous partition
and that round's sampled edges, whatever order the merges are applied in.  This module recomputes
those partitions from the packed bucket images, so the device's per-round diagnostics (vertices
regrouped outside the giant supernode) can be checked round by round.
"""
import numpy as np

import oracle


def parse_images(dump: np.ndarray, n: int, nb: int, gamma_bytes: int = 4):
    """Packed reference images (record i of sketch v at (v*nb + i) * (8 + gamma_bytes)) -> arrays."""
    g = '<u4' if gamma_bytes == 4 else '<u8'
    rec = np.frombuffer(dump.tobytes(), dtype=np.dtype([('a', '<u8'), ('g', g)])).reshape(n, nb)
    return rec['a'].copy(), rec['g'].astype(np.uint64)


def _good(alpha: np.ndarray, gamma: np.ndarray, seed: int, gamma_bytes: int) -> np.ndarray:
    h = oracle.xxh3_len8_many(alpha.astype(np.uint64), np.full(alpha.shape, seed, dtype=np.uint64))
    mask = np.uint64(0xFFFFFFFF) if gamma_bytes == 4 else np.uint64(0xFFFFFFFFFFFFFFFF)
    return (h & mask) == gamma


def sample(alpha_sum: np.ndarray, gamma_sum: np.ndarray, col: int, D: int, seed: int, gamma_bytes: int = 4,
           depth_out: list | None = None):
    """Sketch::sample of a merged sketch on column `col` (src/sketch.cpp:94-116): ('Z'|'G'|'F', index).

    depth_out, if given, gets the depth of a good column bucket that was sampled."""
    da, dg = alpha_sum[-1], gamma_sum[-1]
    if da == 0 and dg == 0:
        return 'Z', 0
    if _good(np.array([da]), np.array([dg]), seed, gamma_bytes)[0]:
        return 'G', int(da)
    ca, cg = alpha_sum[col * D:(col + 1) * D], gamma_sum[col * D:(col + 1) * D]
    good = _good(ca, cg, seed, gamma_bytes)
    if good.any():
        d = int(np.argmax(good))
        if depth_out is not None:
            depth_out.append(d)
        return 'G', int(ca[d])
    return 'F', 0


def partitions(alpha: np.ndarray, gamma: np.ndarray, D: int, seed: int, max_rounds: int, gamma_bytes: int = 4,
               depths: list | None = None):
    """Labels (smallest member of each component) after rounds 0, 1, ... until a round changes nothing.

    depths, if given, gets one list per round: the depth of every good column bucket sampled."""
    n = alpha.shape[0]
    parent = np.arange(n)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    out = []
    for r in range(max_rounds):
        roots = np.array([find(v) for v in range(n)])
        order = np.argsort(roots, kind='stable')
        bounds = np.flatnonzero(np.diff(roots[order])) + 1
        modified = False
        edges = []
        round_depths = []
        if depths is not None:
            depths.append(round_depths)
        for seg in np.split(order, bounds):
            code, idx = sample(np.bitwise_xor.reduce(alpha[seg], axis=0), np.bitwise_xor.reduce(gamma[seg], axis=0),
                               r, D, seed, gamma_bytes, round_depths)
            if code == 'F':
                modified = True
            elif code == 'G':
                a, b = idx >> 32, idx & 0xFFFFFFFF
                if a < n and b < n:
                    edges.append((a, b))
                else:
                    modified = True  # an out-of-range index counts as a failed sample
        for a, b in edges:
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
                modified = True
        labels = np.array([find(v) for v in range(n)])
        out.append(labels)
        if not modified:
            break
    return out
