"""ctypes binding of libgz_hip.so (include/gz_hip.h) with the reference's class and method names.

`Sketch` mirrors include/sketch.h, `CCSketchAlg` mirrors include/cc_sketch_alg.h and
`CCAlgConfiguration` mirrors include/cc_alg_configuration.h.  Errors surface as the reference's
exceptions (OutOfSamplesException, UpdateLockedException).  There is no CPU fallback: if the
library is missing or no GPU is usable, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgz_hip.so")

GZ_OK, GZ_ERR_INVALID_ARG, GZ_ERR_OUT_OF_SAMPLES, GZ_ERR_UPDATE_LOCKED = 0, 1, 2, 3
GZ_ERR_OOM, GZ_ERR_DEVICE, GZ_ERR_ADMISSION, GZ_ERR_IO = 4, 5, 6, 7
GOOD, ZERO, FAIL = 0, 1, 2  # SampleResult, include/sketch.h:20-24
CONNECTIVITY, KSPANNINGFORESTS = 0, 1  # QueryCode, include/cc_sketch_alg.h:60-63
INSERT, DELETE = 0, 1


class GZError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[gz status {code}] {msg}")
        self.code = code


class OutOfSamplesException(GZError):
    """include/sketch.h:193-204"""


class UpdateLockedException(GZError):
    """include/cc_sketch_alg.h:25-29"""


class AdmissionError(GZError):
    """Serving: the tenant's footprint does not fit the memory budget."""


class _AlgConfig(ctypes.Structure):
    _fields_ = [
        ("sketches_factor", ctypes.c_double),
        ("batch_factor", ctypes.c_double),
        ("gamma_bytes", ctypes.c_int),
        ("eager_dsu", ctypes.c_int),
        ("batched", ctypes.c_int),
        ("tile", ctypes.c_uint32),
        ("flush_threshold", ctypes.c_uint64),
    ]


class _AutotuneReport(ctypes.Structure):
    """gz_autotune_report (include/gz_hip.h)."""
    _fields_ = [
        ("trials", ctypes.c_uint32),
        ("tile", ctypes.c_uint32 * 16),
        ("flush_threshold", ctypes.c_uint64 * 16),
        ("ms", ctypes.c_double * 16),
    ]


class Request(ctypes.Structure):
    """gz_request (include/gz_serving.h)."""
    _fields_ = [
        ("kind", ctypes.c_uint32),
        ("tenant", ctypes.c_uint32),
        ("offset", ctypes.c_uint64),
        ("count", ctypes.c_uint64),
        ("u", ctypes.c_uint32),
        ("v", ctypes.c_uint32),
        ("arrival_s", ctypes.c_double),
    ]


class KernelStat(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double), ("alg_bytes", ctypes.c_double),
                ("items", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AlgProfile(ctypes.Structure):
    _fields_ = [("update", KernelStat), ("boruvka", KernelStat)]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("wall_s", ctypes.c_double),
        ("updates", ctypes.c_uint64),
        ("queries", ctypes.c_uint64),
        ("boruvka_runs", ctypes.c_uint64),
        ("cached_answers", ctypes.c_uint64),
        ("query_groups", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


REQ_UPDATES, REQ_POINT, REQ_CC = 0, 1, 2
REQUEST_DTYPE = np.dtype([("kind", "<u4"), ("tenant", "<u4"), ("offset", "<u8"), ("count", "<u8"), ("u", "<u4"),
                          ("v", "<u4"), ("arrival_s", "<f8")], align=True)

_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t

_SIGS = {
    "gz_last_error": (ctypes.c_char_p, []),
    "gz_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "gz_calc_vector_length": (ctypes.c_uint64, [ctypes.c_uint32]),
    "gz_calc_cc_samples": (_sz, [ctypes.c_uint32, ctypes.c_double]),
    "gz_calc_bkt_per_col": (_sz, [ctypes.c_uint64]),
    "gz_sketch_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint64, _sz, _sz, ctypes.c_int, ctypes.POINTER(_vp)]),
    "gz_sketch_destroy": (ctypes.c_int, [_vp]),
    "gz_sketch_update": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "gz_sketch_update_batch": (ctypes.c_int, [_vp, _vp, _sz]),
    "gz_sketch_update_batch_device_noalloc": (ctypes.c_int, [_vp, _vp, _sz]),
    "gz_sketch_sample": (ctypes.c_int, [_vp, _u64p, ctypes.POINTER(ctypes.c_int)]),
    "gz_sketch_merge": (ctypes.c_int, [_vp, _vp]),
    "gz_sketch_range_merge": (ctypes.c_int, [_vp, _vp, _sz, _sz]),
    "gz_sketch_merge_raw_bucket_buffer": (ctypes.c_int, [_vp, _vp]),
    "gz_sketch_zero_contents": (ctypes.c_int, [_vp]),
    "gz_sketch_reset_sample_state": (ctypes.c_int, [_vp]),
    "gz_sketch_serialize": (ctypes.c_int, [_vp, _vp, _sz]),
    "gz_sketch_deserialize": (ctypes.c_int, [_vp, _vp, _sz]),
    "gz_sketch_equal": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_int)]),
    "gz_sketch_bucket_array_bytes": (_sz, [_vp]),
    "gz_sketch_get_buckets": (_sz, [_vp]),
    "gz_sketch_get_columns": (_sz, [_vp]),
    "gz_sketch_get_num_samples": (_sz, [_vp]),
    "gz_sketch_get_bkt_per_col": (_sz, [_vp]),
    "gz_sketch_get_seed": (ctypes.c_uint64, [_vp]),
    "gz_alg_config_default": (None, [ctypes.POINTER(_AlgConfig)]),
    "gz_alg_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(_AlgConfig), ctypes.POINTER(_vp)]),
    "gz_alg_destroy": (ctypes.c_int, [_vp]),
    "gz_alg_allocate_worker_memory": (ctypes.c_int, [_vp, _sz]),
    "gz_alg_get_desired_updates_per_batch": (_sz, [_vp]),
    "gz_alg_get_num_vertices": (ctypes.c_uint32, [_vp]),
    "gz_alg_get_seed": (ctypes.c_uint64, [_vp]),
    "gz_alg_max_rounds": (_sz, [_vp]),
    "gz_alg_sketch_bytes": (_sz, [_vp]),
    "gz_alg_pre_insert": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int]),
    "gz_alg_apply_update_batch": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint32, _vp, _sz]),
    "gz_alg_apply_raw_buckets_update": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp]),
    "gz_alg_update": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]),
    "gz_alg_ingest_pairs_device": (ctypes.c_int, [_vp, _vp, _sz]),
    "gz_alg_ingest_pairs": (ctypes.c_int, [_vp, _vp, _sz]),
    "gz_alg_has_cached_query": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "gz_alg_prep_query": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gz_alg_connected_components": (ctypes.c_int, [_vp, _vp, _u32p]),
    "gz_alg_point_query": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]),
    "gz_alg_point_queries": (ctypes.c_int, [_vp, _vp, _vp, _sz, _vp]),
    "gz_alg_calc_spanning_forest": (ctypes.c_int, [_vp, _vp, _sz, ctypes.POINTER(_sz)]),
    "gz_alg_write_binary": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "gz_alg_construct_from_serialized_data": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_AlgConfig), ctypes.POINTER(_vp)]),
    "gz_alg_dump_sketches": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, _sz]),
    "gz_alg_last_query_rounds": (_sz, [_vp]),
    "gz_alg_get_total_updates": (ctypes.c_uint64, [_vp]),
    "gz_alg_footprint_bytes": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.POINTER(_AlgConfig)]),
    "gz_alg_digest": (ctypes.c_int, [_vp, _u64p]),
    "gz_alg_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gz_alg_set_query_graphs": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gz_alg_set_round_finish": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gz_alg_set_deep_skip": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gz_autotune": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64, _vp, _vp, _sz, _vp, _vp]),
    "gz_alg_last_query_finish_round": (ctypes.c_uint32, [_vp]),
    "gz_alg_last_query_round_sizes": (_sz, [_vp, _vp, _vp, _sz]),
    "gz_alg_get_profile": (ctypes.c_int, [_vp, ctypes.POINTER(AlgProfile)]),
    # include/gz_serving.h
    "gz_alg_serve_trace": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _sz, _vp, _vp, ctypes.POINTER(ServeStats)]),
    "gz_alg_plain_trace": (ctypes.c_int, [_vp, _vp, _vp, _sz, _vp, _vp, ctypes.POINTER(ServeStats)]),
    "gz_server_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_vp)]),
    "gz_server_destroy": (ctypes.c_int, [_vp]),
    "gz_server_admit": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(_AlgConfig),
                                       ctypes.POINTER(ctypes.c_int)]),
    "gz_server_release": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gz_server_tenant_alg": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "gz_server_reserved_bytes": (ctypes.c_uint64, [_vp]),
    "gz_server_budget_bytes": (ctypes.c_uint64, [_vp]),
    "gz_server_run_trace": (ctypes.c_int, [_vp, _vp, _sz, _vp, _vp, _vp, _sz, ctypes.POINTER(_AlgConfig), ctypes.c_int,
                                           _vp, _vp, _vp, _vp, ctypes.POINTER(ServeStats)]),
}


def exported_symbols() -> list[str]:
    return list(_SIGS)


def lib():
    """Load libgz_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m graphzeppelin_amd.build`")
        # torch bundles its own HIP runtime with the same soname (libamdhip64.so.7).  Loading torch
        # first lets the library bind to that copy, so one process never holds two HIP runtimes
        # (a second runtime cannot open the device).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _register_extra(L)
        _lib = L
    return _lib


_EXTRA_SIGS: dict = {}


def _register_extra(L):
    for name, (res, args) in _EXTRA_SIGS.items():
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args


def _check(rc: int):
    if rc == GZ_OK:
        return
    msg = lib().gz_last_error().decode(errors="replace")
    if rc == GZ_ERR_OUT_OF_SAMPLES:
        raise OutOfSamplesException(rc, msg)
    if rc == GZ_ERR_UPDATE_LOCKED:
        raise UpdateLockedException(rc, msg)
    if rc == GZ_ERR_ADMISSION:
        raise AdmissionError(rc, msg)
    raise GZError(rc, msg)


def device_count() -> int:
    c = ctypes.c_int(0)
    _check(lib().gz_device_count(ctypes.byref(c)))
    return c.value


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---- geometry (include/sketch.h:62-76, 181) ---------------------------------------------------
def calc_vector_length(n: int) -> int:
    return int(lib().gz_calc_vector_length(n))


def calc_cc_samples(n: int, f: float = 1.0) -> int:
    return int(lib().gz_calc_cc_samples(n, f))


def calc_bkt_per_col(vector_len: int) -> int:
    return int(lib().gz_calc_bkt_per_col(vector_len))


class Sketch:
    """include/sketch.h:40-191, buckets resident in HBM."""

    def __init__(self, vector_len: int, seed: int, num_samples: int = 1, cols_per_sample: int = 1,
                 gamma_bytes: int = 4):
        h = ctypes.c_void_p()
        _check(lib().gz_sketch_create(vector_len, seed, num_samples, cols_per_sample, gamma_bytes, ctypes.byref(h)))
        self._h = h
        self.gamma_bytes = gamma_bytes

    def close(self):
        if getattr(self, "_h", None):
            lib().gz_sketch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, idx: int):
        _check(lib().gz_sketch_update(self._h, int(idx)))

    def update_batch(self, idx):
        a = np.ascontiguousarray(np.asarray(idx, dtype=np.uint64))
        _check(lib().gz_sketch_update_batch(self._h, _ptr(a), a.size))

    def update_batch_device_noalloc(self, idx_device_ptr: int, count: int):
        """update_batch over `count` uint64 indices in device memory, allocating nothing (at most
        2^26 indices per call; calls that overlap on one GPU race)."""
        _check(lib().gz_sketch_update_batch_device_noalloc(self._h, ctypes.c_void_p(idx_device_ptr), count))

    def sample(self):
        i = ctypes.c_uint64()
        r = ctypes.c_int()
        _check(lib().gz_sketch_sample(self._h, ctypes.byref(i), ctypes.byref(r)))
        return int(i.value), int(r.value)

    def merge(self, other: "Sketch"):
        _check(lib().gz_sketch_merge(self._h, other._h))

    def range_merge(self, other: "Sketch", start_sample: int, n_samples: int):
        _check(lib().gz_sketch_range_merge(self._h, other._h, start_sample, n_samples))

    def merge_raw_bucket_buffer(self, raw: bytes | np.ndarray):
        buf = np.frombuffer(bytes(raw), dtype=np.uint8) if not isinstance(raw, np.ndarray) else np.ascontiguousarray(raw)
        if buf.nbytes < self.bucket_array_bytes():
            raise ValueError("raw bucket buffer too small")
        _check(lib().gz_sketch_merge_raw_bucket_buffer(self._h, _ptr(buf)))

    def zero_contents(self):
        _check(lib().gz_sketch_zero_contents(self._h))

    def reset_sample_state(self):
        _check(lib().gz_sketch_reset_sample_state(self._h))

    def serialize(self) -> bytes:
        out = np.empty(self.bucket_array_bytes(), dtype=np.uint8)
        _check(lib().gz_sketch_serialize(self._h, _ptr(out), out.nbytes))
        return out.tobytes()

    def deserialize(self, data: bytes):
        buf = np.frombuffer(data, dtype=np.uint8)
        _check(lib().gz_sketch_deserialize(self._h, _ptr(buf), buf.nbytes))

    def __eq__(self, other):
        if not isinstance(other, Sketch):
            return NotImplemented
        e = ctypes.c_int()
        _check(lib().gz_sketch_equal(self._h, other._h, ctypes.byref(e)))
        return bool(e.value)

    def bucket_array_bytes(self) -> int:
        return int(lib().gz_sketch_bucket_array_bytes(self._h))

    def get_buckets(self) -> int:
        return int(lib().gz_sketch_get_buckets(self._h))

    def get_columns(self) -> int:
        return int(lib().gz_sketch_get_columns(self._h))

    def get_num_samples(self) -> int:
        return int(lib().gz_sketch_get_num_samples(self._h))

    def get_bkt_per_col(self) -> int:
        return int(lib().gz_sketch_get_bkt_per_col(self._h))

    def get_seed(self) -> int:
        return int(lib().gz_sketch_get_seed(self._h))

    def column_seed(self, c: int) -> int:  # include/sketch.h:175
        return (self.get_seed() + 5 * c) & 0xFFFFFFFFFFFFFFFF

    def checksum_seed(self) -> int:  # include/sketch.h:176
        return self.get_seed()


@dataclass
class CCAlgConfiguration:
    """include/cc_alg_configuration.h plus the device-form knobs."""
    sketches_factor: float = 1.0
    batch_factor: float = 1.0
    gamma_bytes: int = 4
    eager_dsu: bool = True
    batched: bool = True
    tile: int = 0
    flush_threshold: int = 0

    def to_c(self) -> _AlgConfig:
        c = _AlgConfig()
        lib().gz_alg_config_default(ctypes.byref(c))
        c.sketches_factor = self.sketches_factor
        c.batch_factor = self.batch_factor
        c.gamma_bytes = self.gamma_bytes
        c.eager_dsu = 1 if self.eager_dsu else 0
        c.batched = 1 if self.batched else 0
        c.tile = self.tile
        c.flush_threshold = self.flush_threshold
        return c


def autotune(n: int, seed: int, pairs_device_ptr: int, count: int, config: CCAlgConfiguration | None = None):
    """Tile size and flush threshold for this stream shape (gz_autotune): times the device ingest of
    `count` (src,dst) pairs at `pairs_device_ptr` on scratch instances.  Returns the configuration
    (the given one with the fastest pair) and the trials as (tile, flush_threshold, ms)."""
    base = (config or CCAlgConfiguration()).to_c()
    out = _AlgConfig()
    rep = _AutotuneReport()
    _check(lib().gz_autotune(n, seed, ctypes.byref(base), ctypes.c_void_p(pairs_device_ptr), count, ctypes.byref(out),
                             ctypes.byref(rep)))
    cfg = CCAlgConfiguration(out.sketches_factor, out.batch_factor, out.gamma_bytes, bool(out.eager_dsu),
                             bool(out.batched), int(out.tile), int(out.flush_threshold))
    trials = [(int(rep.tile[i]), int(rep.flush_threshold[i]), float(rep.ms[i])) for i in range(rep.trials)]
    return cfg, trials


def footprint_bytes(n: int, config: CCAlgConfiguration | None = None) -> int:
    c = (config or CCAlgConfiguration()).to_c()
    return int(lib().gz_alg_footprint_bytes(n, ctypes.byref(c)))


class ConnectedComponents:
    """include/return_types.h:13-27: parent array of DSU roots and the component count."""

    def __init__(self, parents: np.ndarray, num_cc: int):
        self.parent_arr = parents
        self.num_cc = num_cc

    def size(self) -> int:
        return self.num_cc

    def is_connected(self, a: int, b: int) -> bool:
        return bool(self.parent_arr[a] == self.parent_arr[b])

    def get_component_sets(self) -> list[set[int]]:
        order = np.argsort(self.parent_arr, kind="stable")
        roots = self.parent_arr[order]
        cuts = np.flatnonzero(np.diff(roots)) + 1
        return [set(map(int, grp)) for grp in np.split(order, cuts)]

    def canonical_labels(self) -> np.ndarray:
        """Smallest vertex id of each vertex's component (root-independent form)."""
        lab = np.full(self.parent_arr.size, np.iinfo(np.uint32).max, dtype=np.uint32)
        np.minimum.at(lab, self.parent_arr, np.arange(self.parent_arr.size, dtype=np.uint32))
        return lab[self.parent_arr]


class CCSketchAlg:
    """include/cc_sketch_alg.h:67-257 with every vertex sketch resident in HBM."""

    def __init__(self, num_vertices: int, seed: int, config: CCAlgConfiguration | None = None, _handle=None):
        self.config = config or CCAlgConfiguration()
        if _handle is not None:
            self._h = _handle
        else:
            h = ctypes.c_void_p()
            c = self.config.to_c()
            _check(lib().gz_alg_create(num_vertices, seed, ctypes.byref(c), ctypes.byref(h)))
            self._h = h
        self.num_vertices = int(lib().gz_alg_get_num_vertices(self._h))

    @classmethod
    def construct_from_serialized_data(cls, path: str, config: CCAlgConfiguration | None = None):
        config = config or CCAlgConfiguration()
        h = ctypes.c_void_p()
        c = config.to_c()
        _check(lib().gz_alg_construct_from_serialized_data(path.encode(), ctypes.byref(c), ctypes.byref(h)))
        return cls(0, 0, config, _handle=h)

    def close(self):
        if getattr(self, "_h", None):
            lib().gz_alg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # plugin interface (include/graph_sketch_driver.h:23-69)
    def allocate_worker_memory(self, num_workers: int):
        _check(lib().gz_alg_allocate_worker_memory(self._h, num_workers))

    def get_desired_updates_per_batch(self) -> int:
        return int(lib().gz_alg_get_desired_updates_per_batch(self._h))

    def get_num_vertices(self) -> int:
        return self.num_vertices

    def get_seed(self) -> int:
        return int(lib().gz_alg_get_seed(self._h))

    def max_rounds(self) -> int:
        return int(lib().gz_alg_max_rounds(self._h))

    def sketch_bytes(self) -> int:
        return int(lib().gz_alg_sketch_bytes(self._h))

    def pre_insert(self, src: int, dst: int, type_: int = INSERT, thr_id: int = 0):
        _check(lib().gz_alg_pre_insert(self._h, src, dst, type_, thr_id))

    def apply_update_batch(self, thr_id: int, src: int, dsts):
        a = np.ascontiguousarray(np.asarray(dsts, dtype=np.uint32))
        _check(lib().gz_alg_apply_update_batch(self._h, thr_id, src, _ptr(a), a.size))

    def apply_raw_buckets_update(self, src: int, raw):
        buf = np.ascontiguousarray(np.frombuffer(bytes(raw), dtype=np.uint8))
        if buf.nbytes < self.sketch_bytes():
            raise ValueError("raw bucket buffer too small")
        _check(lib().gz_alg_apply_raw_buckets_update(self._h, src, _ptr(buf)))

    def update(self, src: int, dst: int, type_: int = INSERT):
        _check(lib().gz_alg_update(self._h, src, dst, type_))

    def ingest_pairs(self, pairs: np.ndarray):
        """(src,dst) uint32 pairs in host memory, shape (count, 2)."""
        a = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))
        _check(lib().gz_alg_ingest_pairs(self._h, _ptr(a), a.shape[0]))

    def ingest_pairs_device(self, ptr: int, count: int):
        """(src,dst) uint32 pairs already in device memory (e.g. a torch tensor's data_ptr())."""
        _check(lib().gz_alg_ingest_pairs_device(self._h, ctypes.c_void_p(ptr), count))

    def has_cached_query(self, code: int = CONNECTIVITY) -> bool:
        r = ctypes.c_int()
        _check(lib().gz_alg_has_cached_query(self._h, code, ctypes.byref(r)))
        return bool(r.value)

    def prep_query(self, code: int = CONNECTIVITY):
        _check(lib().gz_alg_prep_query(self._h, code))

    def connected_components(self) -> ConnectedComponents:
        par = np.empty(self.num_vertices, dtype=np.uint32)
        ncc = ctypes.c_uint32()
        _check(lib().gz_alg_connected_components(self._h, _ptr(par), ctypes.byref(ncc)))
        return ConnectedComponents(par, int(ncc.value))

    def point_query(self, a: int, b: int) -> bool:
        r = ctypes.c_int()
        _check(lib().gz_alg_point_query(self._h, a, b, ctypes.byref(r)))
        return bool(r.value)

    def point_queries(self, us, vs) -> np.ndarray:
        u = np.ascontiguousarray(np.asarray(us, dtype=np.uint32))
        v = np.ascontiguousarray(np.asarray(vs, dtype=np.uint32))
        out = np.empty(u.size, dtype=np.uint8)
        _check(lib().gz_alg_point_queries(self._h, _ptr(u), _ptr(v), u.size, _ptr(out)))
        return out.astype(bool)

    def calc_spanning_forest(self) -> np.ndarray:
        cap = max(1, self.num_vertices)
        out = np.empty((cap, 2), dtype=np.uint32)
        k = ctypes.c_size_t()
        _check(lib().gz_alg_calc_spanning_forest(self._h, _ptr(out), cap, ctypes.byref(k)))
        return out[: min(k.value, cap)].copy()

    def write_binary(self, path: str):
        _check(lib().gz_alg_write_binary(self._h, path.encode()))

    def dump_sketches(self, v0: int = 0, count: int | None = None) -> np.ndarray:
        if count is None:
            count = self.num_vertices - v0
        sb = self.sketch_bytes()
        out = np.empty(count * sb, dtype=np.uint8)
        _check(lib().gz_alg_dump_sketches(self._h, v0, count, _ptr(out), out.nbytes))
        return out

    @property
    def last_query_rounds(self) -> int:
        return int(lib().gz_alg_last_query_rounds(self._h))

    def get_total_updates(self) -> int:
        return int(lib().gz_alg_get_total_updates(self._h))

    def set_profiling(self, enable: bool = True):
        _check(lib().gz_alg_set_profiling(self._h, 1 if enable else 0))

    def last_query_round_sizes(self):
        """[(vertices regrouped, giant root or None)] for rounds 1.. of the last Borůvka query."""
        e = np.zeros(64, dtype=np.uint32)
        g = np.zeros(64, dtype=np.uint32)
        k = int(lib().gz_alg_last_query_round_sizes(self._h, e.ctypes.data_as(ctypes.c_void_p),
                                                     g.ctypes.data_as(ctypes.c_void_p), 64))
        return [(int(e[i]), None if g[i] == 0xFFFFFFFF else int(g[i])) for i in range(k)]

    def set_query_graphs(self, enable: bool = True):
        """Replay each query's first Borůvka rounds from a HIP graph (default) or launch them one by one."""
        _check(lib().gz_alg_set_query_graphs(self._h, 1 if enable else 0))

    def set_round_finish(self, enable: bool = True):
        """Run the last Borůvka rounds in one single-workgroup launch (default) or as separate launches."""
        _check(lib().gz_alg_set_round_finish(self._h, 1 if enable else 0))

    def set_deep_skip(self, enable: bool = True):
        """Let Borůvka reduces skip deep records whose column deep flag is clear (default), or read them all."""
        _check(lib().gz_alg_set_deep_skip(self._h, 1 if enable else 0))

    @property
    def last_query_finish_round(self) -> int:
        """Round from which the last query ran its remaining rounds in one workgroup (0: it did not)."""
        return int(lib().gz_alg_last_query_finish_round(self._h))

    def get_profile(self) -> dict:
        p = AlgProfile()
        _check(lib().gz_alg_get_profile(self._h, ctypes.byref(p)))
        return {"update": p.update.as_dict(), "boruvka": p.boruvka.as_dict()}

    def digest(self) -> int:
        d = ctypes.c_uint64()
        _check(lib().gz_alg_digest(self._h, ctypes.byref(d)))
        return int(d.value)

    def serve_trace(self, pairs_ptr: int, on_device: bool, reqs: np.ndarray):
        """Batched/served replay (gz_alg_serve_trace).  `reqs` has REQUEST_DTYPE; returns
        (latency seconds, answers, stats dict)."""
        r = np.ascontiguousarray(reqs, dtype=REQUEST_DTYPE)
        lat = np.zeros(r.size, dtype=np.float64)
        ans = np.zeros(r.size, dtype=np.uint32)
        st = ServeStats()
        _check(lib().gz_alg_serve_trace(self._h, ctypes.c_void_p(pairs_ptr), 1 if on_device else 0, _ptr(r), r.size,
                                        _ptr(lat), _ptr(ans), ctypes.byref(st)))
        return lat, ans, st.as_dict()

    def plain_trace(self, pairs: np.ndarray, reqs: np.ndarray):
        """Reference-shaped replay (gz_alg_plain_trace) over host pairs."""
        p = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint32).reshape(-1, 2))
        r = np.ascontiguousarray(reqs, dtype=REQUEST_DTYPE)
        lat = np.zeros(r.size, dtype=np.float64)
        ans = np.zeros(r.size, dtype=np.uint32)
        st = ServeStats()
        _check(lib().gz_alg_plain_trace(self._h, _ptr(p), _ptr(r), r.size, _ptr(lat), _ptr(ans), ctypes.byref(st)))
        return lat, ans, st.as_dict()


def digest_of_image(image: np.ndarray, num_buckets: int, gamma_bytes: int = 4) -> int:
    """The gz_alg_digest formula over a packed byte image (numpy, mod 2^64)."""
    rec = 8 + gamma_bytes
    raw = np.ascontiguousarray(image, dtype=np.uint8).reshape(-1, rec)
    alpha = raw[:, :8].copy().view("<u8").ravel()
    gamma = raw[:, 8:].copy().view("<u4" if gamma_bytes == 4 else "<u8").ravel().astype(np.uint64)
    pos = np.arange(alpha.size, dtype=np.uint64)

    def fmix(k):
        k = k ^ (k >> np.uint64(33))
        k = k * np.uint64(0xFF51AFD7ED558CCD)
        k = k ^ (k >> np.uint64(33))
        k = k * np.uint64(0xC4CEB9FE1A85EC53)
        return k ^ (k >> np.uint64(33))

    with np.errstate(over="ignore"):
        x = alpha ^ fmix(gamma ^ (pos * np.uint64(0x9E3779B97F4A7C15)))
        return int(np.sum(fmix(x + pos), dtype=np.uint64))


class Server:
    """Multi-tenant request loop (include/gz_serving.h)."""

    def __init__(self, hbm_budget_bytes: int = 0, workers: int = 4):
        h = ctypes.c_void_p()
        _check(lib().gz_server_create(hbm_budget_bytes, workers, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().gz_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def budget_bytes(self) -> int:
        return int(lib().gz_server_budget_bytes(self._h))

    @property
    def reserved_bytes(self) -> int:
        return int(lib().gz_server_reserved_bytes(self._h))

    def admit(self, n: int, seed: int, config: CCAlgConfiguration | None = None) -> int:
        t = ctypes.c_int(-1)
        c = (config or CCAlgConfiguration()).to_c()
        _check(lib().gz_server_admit(self._h, n, seed, ctypes.byref(c), ctypes.byref(t)))
        return t.value

    def release(self, tenant: int):
        _check(lib().gz_server_release(self._h, tenant))

    def run_trace(self, reqs: np.ndarray, tenant_ptrs, tenant_n, tenant_seed, config=None, release_when_done=True,
                  want_parents=False):
        """Replays a mixed trace; returns (latencies, answers, digests, stats), plus each tenant's
        final parent array when want_parents."""
        r = np.ascontiguousarray(reqs, dtype=REQUEST_DTYPE)
        T = len(tenant_n)
        ptrs = (ctypes.c_void_p * T)(*[ctypes.c_void_p(int(p)) for p in tenant_ptrs])
        tn = np.ascontiguousarray(tenant_n, dtype=np.uint32)
        ts = np.ascontiguousarray(tenant_seed, dtype=np.uint64)
        lat = np.zeros(r.size, dtype=np.float64)
        ans = np.zeros(r.size, dtype=np.uint32)
        dig = np.zeros(T, dtype=np.uint64)
        par = [np.zeros(int(n), dtype=np.uint32) for n in tn] if want_parents else None
        par_ptrs = (ctypes.c_void_p * T)(*[_ptr(p) for p in par]) if want_parents else None
        st = ServeStats()
        c = (config or CCAlgConfiguration()).to_c()
        _check(lib().gz_server_run_trace(self._h, _ptr(r), r.size, ctypes.cast(ptrs, ctypes.c_void_p), _ptr(tn),
                                         _ptr(ts), T, ctypes.byref(c), 1 if release_when_done else 0, _ptr(lat),
                                         _ptr(ans), _ptr(dig),
                                         ctypes.cast(par_ptrs, ctypes.c_void_p) if want_parents else None,
                                         ctypes.byref(st)))
        if want_parents:
            return lat, ans, dig, st.as_dict(), par
        return lat, ans, dig, st.as_dict()
