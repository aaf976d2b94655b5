"""The core configuration (the reference's TerrainGenerator: final density and the core fill,
no biomes / aquifers / material rules) through the C ABI, against the reference's outputs
(golden digests of cpu_generate_chunk) and the oracle, bit for bit."""
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


def test_core_edge_chunks_match_reference(mct, golden):
    g = mct.FullTerrainGenerator(max_batch_chunks=8, pipeline_slots=2)
    seeds, xz = golden['edge_seed'], golden['edge_xz']
    d, f = g.generate_core(seeds, [tuple(c) for c in xz])
    g.close()
    for i in range(len(xz)):
        assert (digest(d[i]) == golden['core_density'][i]).all(), (int(seeds[i]), tuple(xz[i]))
        assert (digest(f[i]) == golden['core_fill'][i]).all(), (int(seeds[i]), tuple(xz[i]))


def test_core_region_matches_oracle_and_terrain_path(mct, orc):
    """The 8 x 8 C1 region: density bit for bit and fill against the oracle; one output alone
    equals both together; the device form equals the host form."""
    seed, chunks = workloads.c1_region()
    g = mct.FullTerrainGenerator(max_batch_chunks=24, pipeline_slots=2)
    d, f = g.generate_core(seed, chunks)
    od, of = orc.generate_core(seed, chunks)
    assert (d.view(np.uint32) == od.view(np.uint32)).all()
    assert (f == of).all()
    d2, _ = g.generate_core(seed, chunks, fill=False)
    _, f2 = g.generate_core(seed, chunks, density=False)
    assert (d2.view(np.uint32) == d.view(np.uint32)).all() and (f2 == f).all()
    n = len(chunks)
    dd, df = mct.alloc_device(n * mct.BLOCKS_PER_CHUNK * 4), mct.alloc_device(n * mct.BLOCKS_PER_CHUNK)
    g.generate_core_device(seed, chunks, dd, df)
    assert (mct.device_to_host(dd, n * mct.BLOCKS_PER_CHUNK * 4).view(np.uint32) == d.view(np.uint32).ravel()).all()
    assert (mct.device_to_host(df, n * mct.BLOCKS_PER_CHUNK) == f.ravel()).all()
    mct.free_device(dd)
    mct.free_device(df)
    g.close()


def test_core_rejects_no_output(mct):
    g = mct.FullTerrainGenerator(max_batch_chunks=8, pipeline_slots=1)
    with pytest.raises(mct.MctError) as e:
        g.generate_core(1, [(0, 0)], density=False, fill=False)
    g.close()
    assert e.value.code == 1


def test_topdown_records_match_oracle(mct):
    """Device top-down records of the C1 region (host and device terrain) against the numpy
    restatement of the reference bench's MCTD column walk."""
    import oracle
    seed, chunks = workloads.c1_region()
    g = mct.FullTerrainGenerator(max_batch_chunks=64, pipeline_slots=2)
    g.set_seed(seed)
    t, _ = g.generate(chunks)
    want = oracle.topdown_records(t)
    assert (mct.topdown(t) == want).all()
    n = len(chunks)
    d_t = mct.alloc_device(n * mct.BLOCKS_PER_CHUNK)
    g.generate_device(np.full(n, seed, np.int64), chunks, d_t)
    assert (mct.topdown(d_t, n) == want).all()
    mct.free_device(d_t)
    g.close()


def test_region_dump_tool_end_to_end(mct, tmp_path, monkeypatch):
    """region_dump.py on a 3 x 3 region: the dump holds the generator's bytes in row-major chunk
    order, the MCTD raster the device records placed by chunk."""
    import sys
    import oracle
    import region_dump as rd
    monkeypatch.setattr(sys, 'argv', ['region_dump.py', '--seed', '42', '--x0', '-2', '--z0', '7', '--region', '3',
                                      '--dump', str(tmp_path / 'd'), '--topdown', str(tmp_path / 'v.mctd')])
    rd.main()
    meta, t, b = rd.read_dump(str(tmp_path / 'd'))
    order = rd.region_order(-2, 7, 3)
    g = mct.FullTerrainGenerator(max_batch_chunks=16, pipeline_slots=1)
    g.set_seed(42)
    wt, wb = g.generate(order)
    g.close()
    assert meta['seed'] == 42 and (t == wt).all() and (b == wb).all()
    head, raster = rd.read_topdown(str(tmp_path / 'v.mctd'))
    assert head == {'version': 1, 'region': 3, 'x0': -2, 'z0': 7, 'min_y': -64}
    assert (raster == rd.topdown_raster(oracle.topdown_records(wt), order, -2, 7, 3)).all()
