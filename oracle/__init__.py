"""CPU oracle for GraphZeppelin's streaming connected-components path -- TEST INFRASTRUCTURE ONLY.

A C++ restatement of the reference path (oracle/gz_oracle.cpp) loaded through ctypes.  Only the
tests, the smoke check and the CPU-baseline leg of bench.py use it, always as the checker or the
CPU baseline; the product (graphzeppelin_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgz_oracle.so")

_lib = None
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64

_SIGS = {
    "or_xxh3_len8": (_u64, [_u64, _u64]),
    "or_xxh3_len8_many": (None, [_vp, _vp, _sz, _vp]),
    "or_calc_vector_length": (_u64, [_u32]),
    "or_calc_cc_samples": (_sz, [_u32, ctypes.c_double]),
    "or_calc_bkt_per_col": (_sz, [_u64]),
    "or_sketch_create": (_vp, [_u64, _u64, _sz, _sz, ctypes.c_int]),
    "or_sketch_destroy": (None, [_vp]),
    "or_sketch_update_many": (None, [_vp, _vp, _sz]),
    "or_sketch_sample": (ctypes.c_int, [_vp, ctypes.POINTER(_u64), ctypes.POINTER(ctypes.c_int)]),
    "or_sketch_merge": (None, [_vp, _vp]),
    "or_sketch_range_merge": (None, [_vp, _vp, _sz, _sz]),
    "or_sketch_merge_raw": (None, [_vp, _vp]),
    "or_sketch_zero": (None, [_vp]),
    "or_sketch_reset_sample_state": (None, [_vp]),
    "or_sketch_bytes": (_sz, [_vp]),
    "or_sketch_num_buckets": (_sz, [_vp]),
    "or_sketch_bkt_per_col": (_sz, [_vp]),
    "or_sketch_serialize": (None, [_vp, _vp]),
    "or_alg_create": (_vp, [_u32, _u64, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "or_alg_destroy": (None, [_vp]),
    "or_alg_sketch_bytes": (_sz, [_vp]),
    "or_alg_desired_updates_per_batch": (_sz, [_vp]),
    "or_alg_allocate_worker_memory": (None, [_vp, _sz]),
    "or_alg_apply_update_batch": (ctypes.c_int, [_vp, ctypes.c_int, _u32, _vp, _sz]),
    "or_alg_apply_raw": (None, [_vp, _u32, _vp]),
    "or_alg_pre_insert": (None, [_vp, _u32, _u32]),
    "or_alg_update": (None, [_vp, _u32, _u32]),
    "or_alg_ingest_pairs": (None, [_vp, _vp, _sz, ctypes.c_int]),
    "or_alg_has_cached_query": (ctypes.c_int, [_vp]),
    "or_alg_connected_components": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_u32)]),
    "or_alg_point_query": (ctypes.c_int, [_vp, _u32, _u32, ctypes.POINTER(ctypes.c_int)]),
    "or_alg_last_query_rounds": (_sz, [_vp]),
    "or_alg_spanning_forest": (_sz, [_vp, _vp, _sz]),
    "or_alg_dump": (None, [_vp, _u32, _u32, _vp]),
    "or_exact_cc": (_u32, [_u32, _vp, _sz, _vp]),
    "or_alg_write_binary": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "or_alg_construct_from_serialized": (_vp, [ctypes.c_char_p, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
    "or_baseline_trace": (ctypes.c_int, [_vp, _vp, _sz, _vp, _vp, _vp, _vp, _sz, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_double), _vp, _vp]),
}

GOOD, ZERO, FAIL = 0, 1, 2


class OracleOutOfSamples(RuntimeError):
    pass


def build() -> str:
    """Compile oracle/gz_oracle.cpp with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for k, (r, a) in _SIGS.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def xxh3_len8(idx: int, seed: int) -> int:
    return int(lib().or_xxh3_len8(idx, seed))


def xxh3_len8_many(idx: np.ndarray, seeds: np.ndarray) -> np.ndarray:
    i = np.ascontiguousarray(idx, dtype=np.uint64)
    s = np.ascontiguousarray(np.broadcast_to(seeds, i.shape), dtype=np.uint64)
    out = np.empty_like(i)
    lib().or_xxh3_len8_many(_p(i), _p(s), i.size, _p(out))
    return out


def calc_vector_length(n):
    return int(lib().or_calc_vector_length(n))


def calc_cc_samples(n, f=1.0):
    return int(lib().or_calc_cc_samples(n, f))


def calc_bkt_per_col(v):
    return int(lib().or_calc_bkt_per_col(v))


class Sketch:
    def __init__(self, vector_len, seed, num_samples=1, cols_per_sample=1, gamma_bytes=4):
        self._h = lib().or_sketch_create(vector_len, seed, num_samples, cols_per_sample, gamma_bytes)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_sketch_destroy(self._h)
            self._h = None

    def update(self, idx):
        self.update_batch([idx])

    def update_batch(self, idx):
        a = np.ascontiguousarray(np.asarray(idx, dtype=np.uint64))
        lib().or_sketch_update_many(self._h, _p(a), a.size)

    def sample(self):
        i = _u64()
        r = ctypes.c_int()
        if lib().or_sketch_sample(self._h, ctypes.byref(i), ctypes.byref(r)) == 2:
            raise OracleOutOfSamples()
        return int(i.value), int(r.value)

    def merge(self, o):
        lib().or_sketch_merge(self._h, o._h)

    def range_merge(self, o, start, n):
        lib().or_sketch_range_merge(self._h, o._h, start, n)

    def merge_raw_bucket_buffer(self, raw):
        b = np.frombuffer(bytes(raw), dtype=np.uint8)
        lib().or_sketch_merge_raw(self._h, _p(b))

    def zero_contents(self):
        lib().or_sketch_zero(self._h)

    def reset_sample_state(self):
        lib().or_sketch_reset_sample_state(self._h)

    def bucket_array_bytes(self):
        return int(lib().or_sketch_bytes(self._h))

    def get_buckets(self):
        return int(lib().or_sketch_num_buckets(self._h))

    def get_bkt_per_col(self):
        return int(lib().or_sketch_bkt_per_col(self._h))

    def serialize(self) -> bytes:
        out = np.empty(self.bucket_array_bytes(), dtype=np.uint8)
        lib().or_sketch_serialize(self._h, _p(out))
        return out.tobytes()


class CCAlg:
    def __init__(self, n, seed, sketches_factor=1.0, batch_factor=1.0, gamma_bytes=4, eager=True):
        self.n = n
        self._h = lib().or_alg_create(n, seed, sketches_factor, batch_factor, gamma_bytes, 1 if eager else 0)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_alg_destroy(self._h)
            self._h = None

    def sketch_bytes(self):
        return int(lib().or_alg_sketch_bytes(self._h))

    def get_desired_updates_per_batch(self):
        return int(lib().or_alg_desired_updates_per_batch(self._h))

    def allocate_worker_memory(self, w):
        lib().or_alg_allocate_worker_memory(self._h, w)

    def apply_update_batch(self, thr, src, dsts):
        a = np.ascontiguousarray(np.asarray(dsts, dtype=np.uint32))
        if lib().or_alg_apply_update_batch(self._h, thr, src, _p(a), a.size):
            raise ValueError(f"worker {thr} has no delta sketch: call allocate_worker_memory first")

    def apply_raw_buckets_update(self, src, raw):
        b = np.frombuffer(bytes(raw), dtype=np.uint8)
        lib().or_alg_apply_raw(self._h, src, _p(b))

    def pre_insert(self, a, b):
        lib().or_alg_pre_insert(self._h, a, b)

    def update(self, a, b):
        lib().or_alg_update(self._h, a, b)

    def ingest_pairs(self, pairs, with_pre_insert=True):
        a = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))
        lib().or_alg_ingest_pairs(self._h, _p(a), a.shape[0], 1 if with_pre_insert else 0)

    def has_cached_query(self):
        return bool(lib().or_alg_has_cached_query(self._h))

    def connected_components(self):
        par = np.empty(self.n, dtype=np.uint32)
        c = _u32()
        if lib().or_alg_connected_components(self._h, _p(par), ctypes.byref(c)) == 2:
            raise OracleOutOfSamples()
        return par, int(c.value)

    def point_query(self, a, b):
        r = ctypes.c_int()
        if lib().or_alg_point_query(self._h, a, b, ctypes.byref(r)) == 2:
            raise OracleOutOfSamples()
        return bool(r.value)

    @property
    def last_query_rounds(self):
        return int(lib().or_alg_last_query_rounds(self._h))

    def spanning_forest(self):
        cap = self.n
        out = np.empty((cap, 2), dtype=np.uint32)
        k = int(lib().or_alg_spanning_forest(self._h, _p(out), cap))
        return out[: min(k, cap)].copy()

    def dump(self, v0=0, count=None) -> np.ndarray:
        if count is None:
            count = self.n - v0
        out = np.empty(count * self.sketch_bytes(), dtype=np.uint8)
        lib().or_alg_dump(self._h, v0, count, _p(out))
        return out

    def write_binary(self, path: str):
        """CCSketchAlg::write_binary (src/cc_sketch_alg.cpp:614-623)."""
        if lib().or_alg_write_binary(self._h, path.encode()):
            raise OSError(f"cannot write {path}")

    @classmethod
    def construct_from_serialized_data(cls, path: str, batch_factor=1.0, gamma_bytes=4, eager=True):
        """CCSketchAlg::construct_from_serialized_data (src/cc_sketch_alg.cpp:30-64)."""
        h = lib().or_alg_construct_from_serialized(path.encode(), batch_factor, gamma_bytes, 1 if eager else 0)
        if not h:
            raise OSError(f"cannot read {path}")
        obj = cls.__new__(cls)
        obj._h = h
        with open(path, "rb") as f:
            obj.n = int.from_bytes(f.read(12)[8:12], "little")
        return obj

    def baseline_trace(self, pairs, qpos, kinds, qa, qb, nthreads):
        """Reference-style CPU run over a request trace; returns (wall_s, latencies, answers)."""
        p = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))
        qp = np.ascontiguousarray(qpos, dtype=np.uint64)
        kd = np.ascontiguousarray(kinds, dtype=np.int32)
        a = np.ascontiguousarray(qa, dtype=np.uint32)
        b = np.ascontiguousarray(qb, dtype=np.uint32)
        lat = np.zeros(qp.size, dtype=np.float64)
        ans = np.zeros(qp.size, dtype=np.uint32)
        wall = ctypes.c_double()
        rc = lib().or_baseline_trace(self._h, _p(p), p.shape[0], _p(qp), _p(kd), _p(a), _p(b), qp.size, nthreads,
                                     ctypes.byref(wall), _p(lat), _p(ans))
        if rc == 2:
            raise OracleOutOfSamples()
        return wall.value, lat, ans


def exact_cc(n: int, pairs) -> tuple[int, np.ndarray]:
    """Components of the final edge-parity set; labels are the smallest vertex of each component."""
    p = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))
    lab = np.empty(n, dtype=np.uint32)
    c = int(lib().or_exact_cc(n, _p(p), p.shape[0], _p(lab)))
    return c, lab


def canonical(parents: np.ndarray) -> np.ndarray:
    """Map a parent/root array to smallest-member labels (root-choice independent)."""
    parents = np.asarray(parents, dtype=np.int64)
    lab = np.full(parents.size, np.iinfo(np.int64).max, dtype=np.int64)
    np.minimum.at(lab, parents, np.arange(parents.size, dtype=np.int64))
    return lab[parents].astype(np.uint32)
