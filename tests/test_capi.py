"""CPU checks of the C ABI: the library loads, exports every symbol include/*.h declares, and its
host-only geometry functions agree with the oracle.  No compute call needs a GPU here."""
import ctypes
import os
import re

import pytest

import graphzeppelin_amd as gz
from graphzeppelin_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header: str) -> list[str]:
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ \*]*?\b(gz_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _headers():
    return [h for h in sorted(os.listdir(os.path.join(REPO, "include"))) if h.endswith(".h")]


@pytest.mark.parametrize("header", _headers())
def test_library_exports_every_declared_symbol(header):
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared(header)
    assert names, f"no declarations parsed from {header}"
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"{header}: missing exports {missing}"


def test_binding_covers_the_header():
    names = set(_declared("gz_hip.h"))
    names |= set(_declared("gz_serving.h"))
    assert names <= set(_lib.exported_symbols()) | set(_lib._EXTRA_SIGS)


@pytest.mark.parametrize("n", [2, 3, 100, 1024, 1 << 15, 1 << 18, 1 << 20, (1 << 31)])
def test_host_geometry_matches_oracle(oracle_mod, n):
    assert gz.calc_vector_length(n) == oracle_mod.calc_vector_length(n)
    assert gz.calc_cc_samples(n, 1.0) == oracle_mod.calc_cc_samples(n, 1.0)
    assert gz.calc_cc_samples(n, 2.5) == oracle_mod.calc_cc_samples(n, 2.5)
    vl = gz.calc_vector_length(n)
    assert gz.calc_bkt_per_col(vl) == oracle_mod.calc_bkt_per_col(vl)


def test_footprint_grows_with_n():
    a = gz.footprint_bytes(1 << 12)
    b = gz.footprint_bytes(1 << 18)
    assert 0 < a < b
    # the sketch slab dominates: 2^18 vertices x 901 buckets x 12 bytes (plus each column's padding to a cache line)
    assert b > (1 << 18) * 901 * 12


def test_no_gpu_means_loud_failure():
    """Without a usable device the product raises instead of falling back to the CPU."""
    if gz.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(gz.GZError) as e:
        gz.CCSketchAlg(1024, 42)
    assert e.value.code == _lib.GZ_ERR_DEVICE
    with pytest.raises(gz.GZError):
        gz.Sketch(100, 1)
