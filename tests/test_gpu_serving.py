"""Serving and batching through the C ABI: requests submitted while others are on the GPU are
batched together (continuous batching), admission follows the HBM result arena, and every
result is the same bytes as the request generated alone."""
import threading
import time

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def mct():
    import mct as m
    m.lib()
    return m


def alone(mct, reqs):
    g = mct.FullTerrainGenerator(max_batch_chunks=512, pipeline_slots=2)
    out = []
    for _, seed, chunks in reqs:
        g.set_seed(seed)
        out.append(g.generate(chunks))
    g.close()
    return out


def serve(mct, reqs, budget, max_batch=256, paced=True):
    srv = mct.Server(max_batch_chunks=max_batch, seed_capacity=16, hbm_budget_bytes=budget)
    bufs = [(np.zeros((len(c), 384, 16, 16), np.uint8), np.zeros((len(c), 96, 4, 4), np.uint8)) for _, _, c in reqs]
    ids = {}

    def client():
        t0 = time.perf_counter()
        for k, (t, seed, chunks) in enumerate(reqs):
            if paced:
                dt = t - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
            ids[srv.submit(seed, chunks, bufs[k][0], bufs[k][1])] = k

    th = threading.Thread(target=client)
    th.start()
    th.join()
    srv.drain()
    done = srv.poll()
    st = srv.stats()
    srv.close()
    return bufs, ids, done, st


def small_requests(n, seed_pool=6, max_side=6, rate=400.0):
    pool = workloads.seed_pool(seed_pool)
    reqs = workloads.requests(rate, 10.0, pool=pool, max_side=max_side)
    return reqs[:n]


def test_served_results_equal_each_request_alone(mct):
    # submitted back to back, so that requests queue behind the first batch and are batched
    # together whatever the GPU's speed (paced arrivals can each finish before the next one)
    reqs = small_requests(24)
    bufs, ids, done, st = serve(mct, reqs, budget=24 << 30, paced=False)
    assert st['completed'] == len(reqs) == len(done)
    ref = alone(mct, reqs)
    for rid, lat in done:
        k = ids[rid]
        assert lat > 0
        assert np.array_equal(bufs[k][0], ref[k][0]) and np.array_equal(bufs[k][1], ref[k][1]), k
    assert st['mixed_batches'] > 0, st     # requests were batched together


def test_small_arena_admits_in_turn_and_stays_exact(mct):
    """An arena of a few hundred chunk slots: requests wait for admission, the ring wraps."""
    reqs = small_requests(30, max_side=8, rate=4000.0)
    import mct as m
    probe = m.Server(max_batch_chunks=64, seed_capacity=16, hbm_budget_bytes=64 << 30)
    fixed = (64 << 30) - probe.stats()['arena_chunks'] * (98304 + 1536)
    probe.close()
    budget = fixed + 150 * (98304 + 1536)
    bufs, ids, done, st = serve(mct, reqs, budget=budget, max_batch=64, paced=False)
    assert st['arena_chunks'] < sum(len(c) for _, _, c in reqs)
    assert st['completed'] == len(reqs)
    ref = alone(mct, reqs)
    for rid, _ in done:
        k = ids[rid]
        assert np.array_equal(bufs[k][0], ref[k][0]) and np.array_equal(bufs[k][1], ref[k][1]), k


def test_oversized_request_is_refused(mct):
    import mct as m
    probe = m.Server(max_batch_chunks=16, seed_capacity=4, hbm_budget_bytes=64 << 30)
    fixed = (64 << 30) - probe.stats()['arena_chunks'] * (98304 + 1536)
    probe.close()
    srv = m.Server(max_batch_chunks=16, seed_capacity=4, hbm_budget_bytes=fixed + 10 * (98304 + 1536))
    t = np.zeros((20, 384, 16, 16), np.uint8)
    with pytest.raises(m.MctError) as e:
        srv.submit(1, workloads.region(0, 0, 5, 4), t)
    assert e.value.code == 4
    srv.close()


def test_batched_mixed_seed_workload_sample_matches_oracle(mct, orc):
    """The 256-seed uneven batch in one call; a sample of its chunks against the oracle."""
    work = workloads.batched()
    seeds = [s for s, ch in work for _ in ch]
    chunks = [c for _, ch in work for c in ch]
    g = mct.FullTerrainGenerator(max_batch_chunks=4096, pipeline_slots=2, seed_capacity=256)
    t, b = g.generate_multi(seeds, chunks)
    g.close()
    rng = np.random.default_rng(7)
    for i in rng.choice(len(chunks), 12, replace=False):
        ot, ob, _ = orc.generate(seeds[i], [chunks[i]])
        assert np.array_equal(t[i], ot[0]) and np.array_equal(b[i], ob[0]), (seeds[i], chunks[i])
