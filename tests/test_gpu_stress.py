"""The 1,024-seed stress list and biome-only block queries through the C ABI, against the
reference cpu/ port's outputs (golden digests and answers); the whole 2^20-chunk list in two
different batchings (per-chunk device digests equal); the column-cached zoom of the surface
pass against the plain FP64 zoom on the device."""
import hashlib
import os

import numpy as np
import pytest

import workloads
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def digest(a, size=8):
    return np.frombuffer(hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=size).digest(), np.uint8)


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope='module')
def mct():
    import mct as m
    m.lib()
    return m


@pytest.fixture(scope='module')
def stress_golden():
    with np.load(os.path.join(GOLDEN, 'stress_golden.npz'), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_stress_sample_matches_reference_digests(mct, stress_golden):
    """1,536 chunks of 1,024 seeds in one call (seed groups beyond the resident capacity)."""
    g = stress_golden
    log('stress sample', len(g['index']))
    gen = mct.FullTerrainGenerator(max_batch_chunks=1024, pipeline_slots=2, seed_capacity=256)
    t, b = gen.generate_multi(g['seed'], [tuple(c) for c in g['xz']])
    gen.close()
    bad = [k for k in range(len(t)) if not ((digest(t[k]) == g['terrain'][k]).all() and (digest(b[k]) == g['biomes'][k]).all())]
    assert not bad, (len(bad), [(int(g['seed'][k]), tuple(g['xz'][k])) for k in bad[:5]])


def test_exact_surface_chunks_match_reference_digests(mct):
    """Chunks whose columns end at heights the speculative surface pass did not assume go
    through the exact pass: selective (only columns whose steep test changed are walked again)
    and with every column refilled and walked again (MCT_TUNE_REPLAY_ALL), both equal to the
    reference's digests."""
    with np.load(os.path.join(GOLDEN, 'exact_golden.npz'), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    n = len(g['index'])
    gen = mct.FullTerrainGenerator(max_batch_chunks=64, pipeline_slots=1, seed_capacity=64)
    for rnd, tune in enumerate(({}, {'surface_replay_all': True})):
        log('exact chunks', n, tune)
        if tune:
            gen.set_tuning(tune)
            assert gen.tuning().get('surface_replay_all')
        t, b = gen.generate_multi(g['seed'], [tuple(c) for c in g['xz']])
        bad = [k for k in range(n) if not ((digest(t[k]) == g['terrain'][k]).all() and (digest(b[k]) == g['biomes'][k]).all())]
        assert not bad, (tune, bad)
        assert gen.biome_tie_stats()['exact_surface_redo'] == (rnd + 1) * n
    gen.close()


def test_device_digest_matches_host_digest(mct):
    gen = mct.FullTerrainGenerator(max_batch_chunks=64, pipeline_slots=1)
    chunks = [(0, 0), (-3, 7), (6128, 0)]
    d_t, d_b = mct.alloc_device(3 * mct.BLOCKS_PER_CHUNK), mct.alloc_device(3 * mct.QUARTS_PER_CHUNK)
    gen.generate_device(np.full(3, 5, np.int64), chunks, d_t, d_b)
    t = mct.device_to_host(d_t, 3 * mct.BLOCKS_PER_CHUNK).reshape(3, -1)
    b = mct.device_to_host(d_b, 3 * mct.QUARTS_PER_CHUNK).reshape(3, -1)
    assert (mct.chunk_digests(d_t, 3, mct.BLOCKS_PER_CHUNK) == mct.digest_host(t)).all()
    assert (mct.chunk_digests(d_b, 3, mct.QUARTS_PER_CHUNK) == mct.digest_host(b)).all()
    t[1, 1000] ^= 1                        # one flipped bit changes the digest
    assert (mct.digest_host(t) != mct.chunk_digests(d_t, 3, mct.BLOCKS_PER_CHUNK)).tolist() == [False, True, False]
    mct.free_device(d_t)
    mct.free_device(d_b)
    gen.close()


def test_full_stress_list_two_batchings_agree_and_match_reference_sample(mct, stress_golden):
    """All 2^20 chunks of the 1,024-seed list, streamed through a 65,536-chunk HBM window:
    once in list order (4,096-chunk batches, all seeds resident), once grouped by seed in
    reverse order (1,000-chunk batches, 64 resident seeds, tables rebuilt as seeds return).
    Per-chunk digests of terrain and biomes agree; the 1,536 sampled chunks equal the
    reference's bytes (golden digests)."""
    seeds, which, xz = workloads.stress()
    cs = seeds[which]
    n, W = len(xz), 1 << 16
    g = stress_golden
    sample = dict(zip(g['index'].tolist(), range(len(g['index']))))
    d_t, d_b = mct.alloc_device(W * mct.BLOCKS_PER_CHUNK), mct.alloc_device(W * mct.QUARTS_PER_CHUNK)

    def run(gen, order):
        dig = np.zeros((n, 2), np.uint64)
        checked = 0
        for off in range(0, n, W):
            sel = order[off:off + W]
            gen.generate_device(cs[sel], xz[sel], d_t, d_b)
            dig[sel, 0] = mct.chunk_digests(d_t, len(sel), mct.BLOCKS_PER_CHUNK)
            dig[sel, 1] = mct.chunk_digests(d_b, len(sel), mct.QUARTS_PER_CHUNK)
            for j, i in enumerate(sel.tolist()):
                k = sample.get(i)
                if k is None:
                    continue
                t = mct.device_to_host(d_t + j * mct.BLOCKS_PER_CHUNK, mct.BLOCKS_PER_CHUNK)
                b = mct.device_to_host(d_b + j * mct.QUARTS_PER_CHUNK, mct.QUARTS_PER_CHUNK)
                assert (digest(t) == g['terrain'][k]).all() and (digest(b) == g['biomes'][k]).all(), (i, int(cs[i]), xz[i].tolist())
                checked += 1
        assert checked == len(sample)
        return dig

    a = mct.FullTerrainGenerator(max_batch_chunks=4096, pipeline_slots=2, seed_capacity=1024)
    log('stress pass 1: list order')
    da = run(a, np.arange(n))
    a.close()
    b = mct.FullTerrainGenerator(max_batch_chunks=1000, pipeline_slots=2, seed_capacity=64)
    log('stress pass 2: grouped by seed, reversed')
    db = run(b, np.argsort(which, kind='stable')[::-1].copy())
    b.close()
    mct.free_device(d_t)
    mct.free_device(d_b)
    bad = np.nonzero((da != db).any(axis=1))[0]
    assert len(bad) == 0, (len(bad), bad[:5].tolist())


def test_biome_queries_match_reference(mct, stress_golden):
    g = stress_golden
    log('queries', len(g['query_xyz']))
    gen = mct.FullTerrainGenerator(max_batch_chunks=64, pipeline_slots=1, seed_capacity=1024)
    out = gen.biome_query(g['query_seed'], g['query_xyz'])
    gen.close()
    bad = np.nonzero(out != g['query_biome'])[0]
    assert len(bad) == 0, (len(bad), g['query_xyz'][bad[:5]].tolist())


def test_biome_query_edges_match_oracle(mct, orc):
    """Blocks above and below the world (quart y clamps), at the world border and around 0;
    and the answers do not depend on how the queries are grouped."""
    xs = [-29999999, -29999984, -1, 0, 3, 29999983, 29999999]
    ys = [-2000, -65, -64, 0, 319, 320, 5000]
    xyz = np.array([[x, y, z] for x in xs for y in ys for z in (xs[0], xs[3], xs[-1])], np.int32)
    seeds = np.where(np.arange(len(xyz)) % 2 == 0, 123456789, -1).astype(np.int64)
    gen = mct.FullTerrainGenerator(max_batch_chunks=64, pipeline_slots=1, seed_capacity=4)
    out = gen.biome_query(seeds, xyz)
    half = gen.biome_query(seeds[1::2], xyz[1::2])
    assert len(gen.biome_query(seeds[:0], xyz[:0])) == 0
    gen.close()
    assert (half == out[1::2]).all()
    assert (out == orc.biome_query(seeds, xyz)).all()


def test_biome_query_rejects_coordinates_beyond_range(mct):
    gen = mct.FullTerrainGenerator(max_batch_chunks=16, pipeline_slots=1)
    with pytest.raises(mct.MctError) as e:
        gen.biome_query([1], [[(1 << 30) + 1, 0, 0]])
    gen.close()
    assert e.value.code == 1


def test_cached_zoom_matches_plain_zoom(mct):
    for seed in (123456789, -1, 0):
        checked, bad, ties = mct.zoom_selftest(seed, 1 << 22)
        log('zoom selftest', seed, checked, bad, 'near ties', ties)
        assert checked >= 1 << 22 and bad == 0 and 0 <= ties < checked // 100
