"""gsr_sort_pairs_noalloc (the allocation-free variant of gsr_sort_pairs, its scratch a fixed
device array) against gsr_sort_pairs through the C ABI: two calls in a row, the second with a
smaller input (the fixed scratch keeps the first call's contents), the key count on the device
and on the host, even and odd pass counts; sizes past the fixed scratch are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sorted_both(gsr, ctx, keys, values, key_bits, count=None):
    n = keys.shape[0]
    out = []
    for noalloc in (False, True):
        kb, vb = gsr.DeviceBuffer(ctx, max(8 * n, 8)), gsr.DeviceBuffer(ctx, max(4 * n, 4))
        cb = gsr.DeviceBuffer(ctx, 4)
        kb.write(keys)
        vb.write(values)
        cnt = None
        if count is not None:
            cb.write(np.array([count], np.uint32))
            cnt = cb.ptr
        gsr.sort_pairs_device(ctx, kb.ptr, vb.ptr, n, key_bits=key_bits, count_ptr=cnt, noalloc=noalloc)
        ctx.synchronize()
        out.append((kb.read(np.empty_like(keys)), vb.read(np.empty_like(values))))
        for b in (kb, vb, cb):
            b.close()
    return out


@pytest.mark.parametrize("key_bits", [45, 32])  # 6 passes (even) and 4 passes
def test_noalloc_sort_equals_sort_pairs(gsr_ctx, key_bits):
    import gsr

    rng = np.random.default_rng(key_bits)
    for call, (n, count) in enumerate([(gsr.NOALLOC_SORT_MAX_KEYS, None), (300_001, 250_000)]):
        print(f"case noalloc sort key_bits={key_bits} call={call} n={n} count={count}", flush=True)
        keys = (rng.integers(0, 1 << 62, n, dtype=np.uint64) & np.uint64((1 << key_bits) - 1))
        keys[: n // 3] &= np.uint64(0xFF)  # many ties: stability matters
        values = np.arange(n, dtype=np.uint32)
        (k0, v0), (k1, v1) = _sorted_both(gsr, gsr_ctx, keys, values, key_bits, count)
        assert np.array_equal(k0, k1) and np.array_equal(v0, v1), f"call {call}: noalloc differs from gsr_sort_pairs"
        m = n if count is None else count
        order = np.argsort(keys[:m], kind="stable")
        assert np.array_equal(k1[:m], keys[:m][order]) and np.array_equal(v1[:m], values[:m][order])
        assert np.array_equal(k1[m:], keys[m:]) and np.array_equal(v1[m:], values[m:])


def test_noalloc_sort_odd_passes_and_limits(gsr_ctx):
    import gsr

    rng = np.random.default_rng(7)
    for n in (70_000, 2_049):  # 3 passes of 13+ bits: the result comes back from the scratch copy
        print(f"case noalloc sort 21 bits n={n}", flush=True)
        keys = rng.integers(0, 1 << 21, n, dtype=np.uint64)
        values = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        (k0, v0), (k1, v1) = _sorted_both(gsr, gsr_ctx, keys, values, 21)
        assert np.array_equal(k0, k1) and np.array_equal(v0, v1)
    with pytest.raises(gsr.GsrError, match="fixed scratch"):
        gsr.sort_pairs_device(gsr_ctx, 1, 1, gsr.NOALLOC_SORT_MAX_KEYS + 1, noalloc=True)
