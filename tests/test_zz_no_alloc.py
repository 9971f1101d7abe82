"""skm_argsort_u32_device_static, the radix sort that allocates nothing (its scratch is a static device
array of the library): through the C ABI, two and three calls in a row (each smaller than the one
before, so a call sees what the last one left in the shared scratch), against the allocating
skm_argsort_u32_device on the same keys and against numpy's stable argsort; a count above the
array's maximum is rejected."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MAX_KEYS = 1 << 20


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def static_sort(skm_mod, r, keys_ptr, n, bits, order_ptr):
    return r.argsort_device_static(keys_ptr, n, bits, order_ptr)


def test_static_sort_matches_the_allocating_sort(skm_mod):
    r = skm_mod.Renderer()
    try:
        for n, bits, seed in ((MAX_KEYS, 32, 1), (300_001, 20, 2), (70_000, 13, 3), (5, 32, 4)):
            say(f"case: static argsort n={n} bits={bits}")
            keys = np.random.default_rng(seed).integers(0, 2**bits, n, dtype=np.uint64).astype(np.uint32)
            src, work = skm_mod.DeviceBuffer(4 * n), skm_mod.DeviceBuffer(4 * n)
            o_alloc, o_static = skm_mod.DeviceBuffer(4 * n), skm_mod.DeviceBuffer(4 * n)
            src.upload(keys)
            work.copy_from(src, 4 * n)
            r.argsort_device(work.ptr.value, n, bits, o_alloc.ptr.value)
            work.copy_from(src, 4 * n)
            assert static_sort(skm_mod, r, work.ptr.value, n, bits, o_static.ptr.value) == skm_mod.SKM_OK, r.error()
            got = o_static.download(np.uint32, n)
            assert np.array_equal(got, o_alloc.download(np.uint32, n))
            assert np.array_equal(got, np.argsort(keys, kind="stable"))
            for b in (src, work, o_alloc, o_static):
                b.close()
        big = skm_mod.DeviceBuffer(4 * (MAX_KEYS + 1))
        rc = static_sort(skm_mod, r, big.ptr.value, MAX_KEYS + 1, 32, big.ptr.value)
        assert rc == skm_mod.SKM_ERR_INVALID and "maximum" in r.error()
        assert static_sort(skm_mod, r, big.ptr.value, 10, 0, big.ptr.value) == skm_mod.SKM_ERR_INVALID
        big.close()
    finally:
        r.close()
