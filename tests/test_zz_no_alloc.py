"""The sketch-batch update that allocates nothing (gz_sketch_update_batch_device_noalloc) against
gz_sketch_update_batch, through the C ABI on an MI355X: byte-identical sketch images after two calls
in a row, the second with a smaller input, at both checksum widths; and the size limit."""
import numpy as np
import pytest

import graphzeppelin_amd as gz
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gb", [4, 8])
def test_noalloc_update_equals_update_batch(gb):
    import torch

    print(f"no-alloc sketch update gb={gb}", flush=True)
    n = 1 << 18
    vl, samples = gz.calc_vector_length(n), gz.calc_cc_samples(n)
    rng = np.random.default_rng(100 + gb)
    a = gz.Sketch(vl, 42, samples, 1, gb)
    b = gz.Sketch(vl, 42, samples, 1, gb)
    r = oracle.Sketch(vl, 42, samples, 1, gb)
    for count in (300_000, 4_321):  # two calls in a row, the second smaller
        idx = rng.integers(0, vl, size=count, dtype=np.uint64)
        dev = torch.from_numpy(idx.view(np.int64)).cuda()
        torch.cuda.synchronize()
        a.update_batch(idx)
        b.update_batch_device_noalloc(dev.data_ptr(), count)
        r.update_batch(idx)
        img = b.serialize()
        assert img == a.serialize(), f"count {count}"
        assert img == r.serialize(), f"count {count}"
    for _ in range(samples):
        assert b.sample() == a.sample()


def test_noalloc_single_tile_and_limit():
    import torch

    vl = gz.calc_vector_length(1024)
    a = gz.Sketch(vl, 7, 18, 1)
    b = gz.Sketch(vl, 7, 18, 1)
    idx = np.array([0, 0, 0, 5, 123456], dtype=np.uint64)  # one tile (no atomic write-back)
    dev = torch.from_numpy(idx.view(np.int64)).cuda()
    torch.cuda.synchronize()
    a.update_batch(idx)
    b.update_batch_device_noalloc(dev.data_ptr(), idx.size)
    assert a.serialize() == b.serialize()
    b.update_batch_device_noalloc(dev.data_ptr(), 0)  # nothing to do
    assert a.serialize() == b.serialize()
    with pytest.raises(gz.GZError) as ei:  # rejected on the host before any launch
        b.update_batch_device_noalloc(dev.data_ptr(), (1 << 26) + 1)
    assert ei.value.code == 1
    assert a.serialize() == b.serialize()
