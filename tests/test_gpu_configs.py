"""Every chunk of the batched and served configurations against the reference cpu/ port's own
per-chunk digests (tests/golden/make_digests.py), compared as 64-bit chunk digests: on the
device (mct_chunk_digests) for outputs kept in HBM, with the same digest on the host
(mct.digest_host) for served results copied to host memory. Also the full plain workload at
the launch shapes the bench picks (64-lane workgroups everywhere)."""
import os
import threading
import time

import numpy as np
import pytest

import workloads
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

PER, QPER = 384 * 16 * 16, 96 * 4 * 4


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope='module')
def mct():
    import mct as m
    m.lib()
    return m


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def device_digests(mct, seeds, xz, **gen_args):
    n = len(xz)
    d_t, d_b = mct.alloc_device(n * PER), mct.alloc_device(n * QPER)
    try:
        g = mct.FullTerrainGenerator(**gen_args)
        g.generate_device(seeds, xz, d_t, d_b)
        g.close()
        return mct.chunk_digests(d_t, n, PER), mct.chunk_digests(d_b, n, QPER)
    finally:
        mct.free_device(d_t)
        mct.free_device(d_b)


def mismatches(got, want):
    return np.flatnonzero(got != want)


@pytest.mark.parametrize('seed_capacity', [256, 48])
def test_batched_workload_every_chunk_matches_reference(mct, seed_capacity):
    """The 256-seed uneven workload (33,966 chunks, half rectangles, half scattered) in one
    mixed-seed call; with 48 resident seed tables the call also evicts and rebuilds tables."""
    f = load('c3_digests.npz')
    log('C3 batched, seed capacity', seed_capacity, 'chunks', len(f['xz']))
    dt, db = device_digests(mct, f['seed'], f['xz'], max_batch_chunks=8192, pipeline_slots=2, seed_capacity=seed_capacity)
    bad_t, bad_b = mismatches(dt, f['terrain']), mismatches(db, f['biomes'])
    assert len(bad_t) == 0 and len(bad_b) == 0, ('terrain', len(bad_t), f['xz'][bad_t[:8]].tolist(), 'biomes', len(bad_b))


def test_served_configuration_every_chunk_matches_reference(mct):
    """The served configuration: Poisson arrivals at the three fixed rates, seeds from the
    64-seed pool, W x H regions (W, H uniform 1..32) in spiral order. The server keeps 16 seed
    tables, so requests of other seeds evict tables while batches are in flight. Every chunk of
    the first SERVE_CHECKED requests of each stream against the reference digests."""
    f = load('c4_digests.npz')
    counts = f['request_chunks']
    off = np.concatenate([[0], np.cumsum(counts)])
    pool = workloads.seed_pool()
    srv = mct.Server(max_batch_chunks=1024, seed_capacity=16, hbm_budget_bytes=48 << 30)
    k = 0
    try:
        for rate in workloads.SERVE_RATES:
            reqs = workloads.requests(rate, 10.0 * workloads.SERVE_CHECKED / rate, pool=pool)[:workloads.SERVE_CHECKED]
            log('C4 served, rate', rate, 'requests', len(reqs), 'chunks', sum(len(c) for _, _, c in reqs))
            bufs = [(np.zeros((len(c), 384, 16, 16), np.uint8), np.zeros((len(c), 96, 4, 4), np.uint8)) for _, _, c in reqs]
            for j, (_, seed, chunks) in enumerate(reqs):
                i = k + j
                assert seed == f['seed'][off[i]] and len(chunks) == counts[i]
                assert np.array_equal(np.array(chunks, np.int32), f['xz'][off[i]:off[i + 1]])

            def client():
                t0 = time.perf_counter()
                for j, (t, seed, chunks) in enumerate(reqs):
                    dt = t - (time.perf_counter() - t0)
                    if dt > 0:
                        time.sleep(dt)
                    srv.submit(seed, chunks, bufs[j][0], bufs[j][1])
            th = threading.Thread(target=client)
            th.start()
            th.join()
            srv.drain()
            done = srv.poll()
            assert len(done) == len(reqs)
            for j in range(len(reqs)):
                i = k + j
                want_t, want_b = f['terrain'][off[i]:off[i + 1]], f['biomes'][off[i]:off[i + 1]]
                bad_t = mismatches(mct.digest_host(bufs[j][0]), want_t)
                bad_b = mismatches(mct.digest_host(bufs[j][1]), want_b)
                assert len(bad_t) == 0 and len(bad_b) == 0, ('rate', rate, 'request', j, 'seed', int(f['seed'][off[i]]), len(bad_t), len(bad_b))
            k += len(reqs)
        st = srv.stats()
        log('C4 server stats', st)
        assert st['completed'] == k
    finally:
        srv.close()


def test_plain_workload_at_64_lane_shapes_matches_reference(mct):
    """Every chunk of the plain three-region workload with every tunable kernel at 64-lane
    workgroups (the shapes the bench's tuning picks for the fill and the one-wave paths)."""
    f = load('plain_digests.npz')
    seed = int(f['seed'][0])
    xz = f['xz']
    g = mct.FullTerrainGenerator(max_batch_chunks=8192, pipeline_slots=2)
    t = g.tuning()
    t['workgroup'] = {k: (64 if v else 0) for k, v in t['workgroup'].items()}
    g.set_tuning(t)
    n = len(xz)
    d_t, d_b = mct.alloc_device(n * PER), mct.alloc_device(n * QPER)
    try:
        g.generate_device(np.full(n, seed, np.int64), xz, d_t, d_b)
        out = np.empty((n, 384, 16, 16), np.uint8)
        bout = np.empty((n, 96, 4, 4), np.uint8)
        out[:] = mct.device_to_host(d_t, n * PER).reshape(out.shape)
        bout[:] = mct.device_to_host(d_b, n * QPER).reshape(bout.shape)
    finally:
        g.close()
        mct.free_device(d_t)
        mct.free_device(d_b)
    import hashlib
    dig = lambda a: np.frombuffer(hashlib.blake2b(np.ascontiguousarray(a).tobytes(), digest_size=8).digest(), np.uint8)
    bad = [i for i in range(n) if not ((dig(out[i]) == f['terrain'][i]).all() and (dig(bout[i]) == f['biomes'][i]).all())]
    assert not bad, ('mismatching chunks', len(bad), xz[bad[:8]].tolist())


def test_edge_set_every_chunk_matches_reference(mct):
    """The edge set: a chunk on each side of the noise wrap fold for every distinct horizontal
    scale x frequency product of the IR (under two seeds), and the reference's 8 x 8 border
    verify region at chunk (1874992, 1874992) (bench/run_all.py:27)."""
    f = load('edge_digests.npz')
    folds, border = workloads.fold_chunks(), workloads.border_region()
    assert f['xz'].tolist() == [list(c) for c in folds + folds + border]
    log('edge set: fold chunks', len(folds), 'x 2 seeds, border region', len(border))
    for batch in (256, 64):
        dt, db = device_digests(mct, f['seed'], f['xz'], max_batch_chunks=batch, pipeline_slots=2, seed_capacity=4)
        bad_t, bad_b = mismatches(dt, f['terrain']), mismatches(db, f['biomes'])
        assert len(bad_t) == 0 and len(bad_b) == 0, ('batch', batch, 'terrain', f['xz'][bad_t[:8]].tolist(), 'biomes', f['xz'][bad_b[:8]].tolist())


def test_stress_every_chunk_and_query_matches_reference(mct):
    """The stress configuration at full size: all 2^20 chunks of the 1,024-seed list (every
    seed's tables resident, in 65,536-chunk HBM windows) and all 2^24 of its biome-only
    queries, against the reference's group digests: each group of 1,024 chunk
    digests (terrain and biomes) and of 4,096 query answers folded with the chunk digest, so a
    group matches when every member matches (up to 64-bit collisions). The first group is also
    compared chunk by chunk, to name the chunks of a mismatch."""
    if not os.path.exists(os.path.join(GOLDEN, 'c5_digests.npz')):
        pytest.skip('tests/golden/c5_digests.npz not written yet (tests/golden/make_digests.py --only c5)')
    f = load('c5_digests.npz')
    G, QG = int(f['group']), int(f['qgroup'])
    seeds, which, xz = workloads.stress()
    n = len(xz)
    assert n == len(f['terrain_groups']) * G
    window = 1 << 16
    dt, db = np.empty(n, np.uint64), np.empty(n, np.uint64)
    d_t, d_b = mct.alloc_device(window * PER), mct.alloc_device(window * QPER)
    g = mct.FullTerrainGenerator(max_batch_chunks=16384, pipeline_slots=2, seed_capacity=1024)
    try:
        for off in range(0, n, window):
            k = min(window, n - off)
            if off % (1 << 18) == 0:
                log('C5 chunks from', off)
            g.generate_device(seeds[which[off:off + k]], xz[off:off + k], d_t, d_b)
            dt[off:off + k] = mct.chunk_digests(d_t, k, PER)
            db[off:off + k] = mct.chunk_digests(d_b, k, QPER)
        head_t, head_b = mismatches(dt[:G], f['head_terrain']), mismatches(db[:G], f['head_biomes'])
        assert len(head_t) == 0 and len(head_b) == 0, ('first group', xz[head_t[:8]].tolist(), xz[head_b[:8]].tolist())
        bad_t = mismatches(mct.digest_host(dt.reshape(-1, G)), f['terrain_groups'])
        bad_b = mismatches(mct.digest_host(db.reshape(-1, G)), f['biome_groups'])
        assert len(bad_t) == 0 and len(bad_b) == 0, ('terrain groups', bad_t[:8].tolist(), 'biome groups', bad_b[:8].tolist())
        qs, q = workloads.queries(int(f['queries']), seeds)
        log('C5 queries', len(q))
        ans = g.biome_query(qs, q)
        assert np.array_equal(ans[:QG], f['head_answers']), np.flatnonzero(ans[:QG] != f['head_answers'])[:8].tolist()
        bad_q = mismatches(mct.digest_host(ans.reshape(-1, QG)), f['query_groups'])
        assert len(bad_q) == 0, ('query groups', len(bad_q), bad_q[:8].tolist())
    finally:
        g.close()
        mct.free_device(d_t)
        mct.free_device(d_b)
