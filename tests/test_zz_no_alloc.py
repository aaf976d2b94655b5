"""The allocation-free digest entry point (mct_chunk_digests_noalloc) through the C ABI against
mct_chunk_digests on the same device data: two calls in a row, the second with fewer chunks
(the fixed buffer keeps the first call's values beyond them), and its capacity error."""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu

PER, QPER = 384 * 16 * 16, 96 * 4 * 4


@pytest.fixture(scope='module')
def mct():
    import mct as m
    m.lib()
    return m


def test_noalloc_digests_equal_allocating_digests(mct):
    chunks = workloads.region(-3, 5, 6, 4) + [(6128, 0), (1874999, -1874999)]
    n = len(chunks)
    seeds = np.array([123456789] * (n - 2) + [1, -1], np.int64)
    d_t, d_b = mct.alloc_device(n * PER), mct.alloc_device(n * QPER)
    try:
        g = mct.FullTerrainGenerator(max_batch_chunks=64, pipeline_slots=2, seed_capacity=4)
        g.generate_device(seeds, chunks, d_t, d_b)
        g.close()
        first_t = mct.chunk_digests_noalloc(d_t, n, PER)
        assert np.array_equal(first_t, mct.chunk_digests(d_t, n, PER))
        second_b = mct.chunk_digests_noalloc(d_b, n - 7, QPER)   # fewer chunks, other data
        assert np.array_equal(second_b, mct.chunk_digests(d_b, n - 7, QPER))
        assert np.array_equal(mct.chunk_digests_noalloc(d_t, 3, PER), first_t[:3])
        assert len(mct.chunk_digests_noalloc(d_t, 0, PER)) == 0
    finally:
        mct.free_device(d_t)
        mct.free_device(d_b)


def test_noalloc_digests_refuse_counts_beyond_the_buffer(mct):
    d = mct.alloc_device(4096)
    try:
        with pytest.raises(mct.MctError) as e:
            mct.chunk_digests_noalloc(d, mct.DIGESTS_MAX + 1, 8)
        assert e.value.code == 4 and 'MCT_DIGESTS_MAX' in str(e.value)
        with pytest.raises(mct.MctError) as e:
            mct.chunk_digests_noalloc(d, 1, 12)
        assert e.value.code == 1
    finally:
        mct.free_device(d)
