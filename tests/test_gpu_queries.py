"""The 1,024-seed stress list and biome-only block queries through the C ABI, against the
reference cpu/ port's outputs (golden digests and answers), plus the column-cached zoom of
the surface pass against the plain FP64 zoom on the device."""
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
        checked, bad, refills = mct.zoom_selftest(seed, 1 << 22)
        log('zoom selftest', seed, checked, bad, refills)
        assert checked >= 1 << 22 and bad == 0
