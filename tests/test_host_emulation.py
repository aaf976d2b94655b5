"""The host-compiled twin of the HIP pipeline (libmct_host.so, the same generated evaluators
and kernel bodies) against the reference's golden digests: the edge chunks, the C1 region,
a stress sample and biome queries. CPU only."""
import hashlib
import os

import numpy as np
import pytest

import workloads
from conftest import GOLDEN

gen_host = pytest.importorskip('gen_host')


def digest(a, size=8):
    return np.frombuffer(hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=size).digest(), np.uint8)


@pytest.fixture(scope='module')
def host():
    if not os.path.exists(gen_host.LIB_PATH):
        pytest.skip('libmct_host.so not built')
    return gen_host.HostGenerator()


def test_edge_chunks(host, golden):
    for i, (s, c) in enumerate(zip(golden['edge_seed'], golden['edge_xz'])):
        t, b = host.generate(int(s), [tuple(c)], threads=1)
        assert (digest(t[0]) == golden['edge_terrain'][i]).all() and (digest(b[0]) == golden['edge_biomes'][i]).all(), (s, c)


def test_c1_region_in_parallel(host, golden):
    seed, chunks = workloads.c1_region()
    t, b = host.generate(seed, chunks, threads=8)
    for i in range(len(chunks)):
        assert (digest(t[i]) == golden['c1_terrain'][i]).all() and (digest(b[i]) == golden['c1_biomes'][i]).all(), i


def test_stress_sample_and_queries(host):
    with np.load(os.path.join(GOLDEN, 'stress_golden.npz'), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    for k in range(0, len(g['index']), 191):
        t, b = host.generate(int(g['seed'][k]), [tuple(g['xz'][k])], threads=1)
        assert (digest(t[0]) == g['terrain'][k]).all() and (digest(b[0]) == g['biomes'][k]).all(), k
    sel = np.arange(0, len(g['query_xyz']), 97)
    for i in sel:
        got = host.biome_query(int(g['query_seed'][i]), g['query_xyz'][i:i + 1])
        assert got[0] == g['query_biome'][i], i


def test_operation_counts_add_over_independent_chunks():
    """The counting build: per-chunk stages (fill, aquifer cells, surface columns) count the
    same operations for a chunk alone as inside a batch, the deduplicated stages add over
    chunks that share no position, and the plain build counts nothing."""
    far = [(0, 0), (5000, -7000)]
    both = gen_host.op_counts(123456789, far)
    one = [gen_host.op_counts(123456789, [c]) for c in far]
    for st in ('fill', 'aquifer_cells', 'surface_columns', 'corners', 'columns', 'fts', 'biome_columns', 'biome_quarts'):
        for c in gen_host.OP_CLASSES:
            assert both[st]['ops'][c] == one[0][st]['ops'][c] + one[1][st]['ops'][c], (st, c)
    assert both['fill']['ops']['f32'] > 0 and both['corners']['ops']['f64'] > 0
    assert gen_host.lib().mcte_counts_ops() == 0


def test_device_division_rewrites_are_exact():
    """The device computes the squeeze's t / 24 and the aquifer pressures' divisions by 25, 1.5,
    2.5, 3 and 10 as a product corrected once by a fused multiply-add, and fuses the noise
    gradient terms g * v + c. Each rewrite equals the IEEE operation bit for bit on its whole
    input domain (every float of the squeeze's range, every half-integer the pressures can
    take); the uncorrected product does not, which shows the check can fail."""
    if not os.path.exists(gen_host.LIB_PATH):
        pytest.skip('libmct_host.so not built')
    r = gen_host.check_divisions(threads=8)
    assert r['squeeze'] == 0
    assert all(v == 0 for v in r['aquifer'].values()), r['aquifer']
    assert all(v > 0 for v in r['raw'].values()), r['raw']
    assert r['gradient'] == 0


def test_java_min_max_clamp_select_forms():
    """Math.min / Math.max (float) and Mth.clamp are computed with selects instead of the JDK's
    early returns; they return the JDK's bits for NaNs of either sign, zeros of either sign,
    infinities and subnormals, on random bit patterns, and for every float under the clamp
    bounds the density functions use."""
    if not os.path.exists(gen_host.LIB_PATH):
        pytest.skip('libmct_host.so not built')
    r = gen_host.check_minmax(threads=8)
    assert r == {'special': 0, 'random': 0, 'clamp': 0}, r


def test_dense_dedupe_index_is_an_order_preserving_bijection():
    """The dense dedupe ranks keys by their index in the codec's dense space: every key of a
    batch maps inside the space, dense_key inverts the map, and distinct keys keep their order
    (so the dense form returns the radix sort's unique keys and map). Random batches of 1 to 4
    clusters and 3 seed slots, for the column, grid and biome codecs."""
    if not os.path.exists(gen_host.LIB_PATH):
        pytest.skip('libmct_host.so not built')
    r = gen_host.check_dense()
    assert r == {'outside': 0, 'roundtrip': 0, 'order': 0, 'batches': 40}, r


def test_exact_surface_chunks(host):
    """Chunks whose columns end at heights the speculative surface pass did not assume: each
    goes through the exact pass and matches the reference's digests, both with the selective
    replay the device runs (only columns whose steep test the live heights change are walked
    again) and with every column of the chunk refilled and walked again in x-major order."""
    with np.load(os.path.join(GOLDEN, 'exact_golden.npz'), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    assert len(g['index']) >= 8
    full = gen_host.HostGenerator(replay_all=True)
    for k in range(len(g['index'])):
        for h, cols in ((host, None), (full, 256)):
            t, b = h.generate(int(g['seed'][k]), [tuple(g['xz'][k])], threads=1)
            assert h.exact_redone == 1, k
            assert cols is None or h.exact_columns == cols, (k, h.exact_columns)
            assert (digest(t[0]) == g['terrain'][k]).all() and (digest(b[0]) == g['biomes'][k]).all(), (k, cols)
