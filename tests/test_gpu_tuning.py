"""Launch shapes and the autotuner: workgroup sizes and quarts per R-tree search thread change
nothing in the results (bit for bit, against the reference's digests), the tuner returns valid
shapes, and bad shapes are refused."""
import hashlib

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


def digest(a, size=8):
    return np.frombuffer(hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=size).digest(), np.uint8)


@pytest.fixture(scope='module')
def mct():
    import mct as m
    m.lib()
    return m


def c1_matches(gen, golden):
    seed, chunks = workloads.c1_region()
    gen.set_seed(seed)
    t, b = gen.generate(chunks)
    return all((digest(t[i]) == golden['c1_terrain'][i]).all() and (digest(b[i]) == golden['c1_biomes'][i]).all() for i in range(len(chunks)))


def test_launch_shapes_do_not_change_results(mct, golden):
    g = mct.FullTerrainGenerator(max_batch_chunks=64, pipeline_slots=2)
    groups = list(g.tuning()['workgroup'])
    edge_seeds, edge_xz = golden['edge_seed'], [tuple(c) for c in golden['edge_xz']]
    base_t, base_b = g.generate_multi(edge_seeds, edge_xz)
    for wg, run_len in ((64, 4), (128, 16), (256, 1), (64, 96)):
        g.set_tuning({'workgroup': {k: wg for k in groups}, 'run_len': run_len})
        assert g.tuning()['run_len'] == run_len
        t, b = g.generate_multi(edge_seeds, edge_xz)
        assert (t == base_t).all() and (b == base_b).all(), (wg, run_len)
        assert c1_matches(g, golden), (wg, run_len)
    g.close()


def test_autotune_returns_valid_shapes_and_stays_exact(mct, golden):
    g = mct.FullTerrainGenerator(max_batch_chunks=256, pipeline_slots=2)
    seed, _ = workloads.c1_region()
    sample = workloads.region(-8, -8, 16)
    tuned = g.autotune(seed, sample)
    print('tuning', tuned, flush=True)
    assert tuned['run_len'] in (4, 6, 8, 12, 16)
    assert tuned['workgroup'] and all(v in (64, 128, 256) for v in tuned['workgroup'].values())
    assert tuned['ms'] and all(v > 0 for v in tuned['ms'].values())
    assert g.tuning()['workgroup'] == tuned['workgroup']
    assert c1_matches(g, golden)
    g.close()


def test_bad_shapes_are_refused(mct):
    g = mct.FullTerrainGenerator(max_batch_chunks=16, pipeline_slots=1)
    groups = list(g.tuning()['workgroup'])
    for bad in ({'workgroup': {groups[0]: 100}, 'run_len': 8}, {'workgroup': {}, 'run_len': 0}, {'workgroup': {}, 'run_len': 97},
                {'workgroup': {'corners': 64}, 'run_len': 8}):
        with pytest.raises(mct.MctError) as e:
            g.set_tuning(bad)
        assert e.value.code == 1, bad
    g.close()
