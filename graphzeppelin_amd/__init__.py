"""MI355X-native streaming connected components (GraphZeppelin's hot path on gfx950).

The compute lives in libgz_hip.so (hand-written HIP kernels behind the C ABI of include/gz_hip.h);
this package is its Python face: a ctypes binding with the reference's class names, stream formats
and generators.
"""
from ._lib import (  # noqa: F401
    CONNECTIVITY,
    FAIL,
    GOOD,
    KSPANNINGFORESTS,
    ZERO,
    AdmissionError,
    CCAlgConfiguration,
    CCSketchAlg,
    ConnectedComponents,
    REQUEST_DTYPE,
    REQ_CC,
    REQ_POINT,
    REQ_UPDATES,
    Server,
    digest_of_image,
    GZError,
    OutOfSamplesException,
    Sketch,
    UpdateLockedException,
    autotune,
    calc_bkt_per_col,
    calc_cc_samples,
    calc_vector_length,
    device_count,
    footprint_bytes,
    lib,
)

__all__ = [
    "CCAlgConfiguration",
    "CCSketchAlg",
    "ConnectedComponents",
    "Sketch",
    "GZError",
    "OutOfSamplesException",
    "UpdateLockedException",
    "AdmissionError",
    "GOOD",
    "ZERO",
    "FAIL",
    "CONNECTIVITY",
    "KSPANNINGFORESTS",
    "calc_vector_length",
    "autotune",
    "calc_cc_samples",
    "calc_bkt_per_col",
    "device_count",
    "footprint_bytes",
    "lib",
]
