"""Writes the full-size reference digests of the batched, served and stress configurations from
the reference cpu/ port compiled from its own sources (`make -C oracle ref`,
oracle/_ref/libmcref.so). Runs in the build container only; what it writes is data.

  python tests/golden/make_digests.py [--threads T] [--only c3,c4,c5q,c5,edge]

Every digest is the library's 64-bit chunk digest (the sum over 8-byte words of a SplitMix64
mix, as `mct_chunk_digests` computes it on the device), so a GPU test compares HBM outputs
with these files without copying the outputs to the host.

  c3_digests.npz   the 256-seed batched workload (workloads.batched(), 33,966 chunks in list
                   order): terrain and biome digest of every chunk
  c4_digests.npz   the served configuration: the first SERVE_CHECKED requests of the Poisson
                   stream at each of the SERVE_RATES (64-seed pool, W x H uniform 1..32,
                   spiral order): terrain and biome digest of every chunk, in request order
  c5_digests.npz   the stress configuration: all 2^20 chunks of the 1,024-seed list and all
                   2^24 biome-only block queries (the reference answers a query by generating
                   its chunk's biomes, about 1,450 s for the 2^24 on 8 cores). Stored as
                   digests of digests: group g of
                   GROUP chunks holds digest(terrain digests of its chunks as 8-byte words),
                   likewise for biomes, and group g of QGROUP queries digest(its answer bytes).
                   A GPU run that folds its own per-chunk digests the same way matches every
                   group exactly when every chunk matches (up to 64-bit collisions).
  edge_digests.npz the edge set: a chunk on each side of the noise wrap fold for every
                   distinct horizontal scale x frequency product of the IR
                   (workloads.fold_chunks(), under the bench seed and under seed 1) and the
                   reference's 8 x 8 border verify region at chunk (1874992, 1874992)

The 2^20-chunk pass takes about 45,000 CPU-seconds; partial results are kept under
--state (default /tmp/mct_digest_state) so an interrupted run resumes.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, 'oracle'))
sys.path.insert(0, os.path.join(ROOT, 'minecraft-terrain-cuda_amd'))
import oracle  # noqa: E402
import workloads  # noqa: E402

GROUP = 1024      # chunks per stress digest group
QGROUP = 4096     # queries per stress query group
QUERIES = 1 << 24   # every query of the stress configuration (the reference answers each with
                    # a whole chunk's biomes: about 1,600 s on 8 cores)


def fold(words):
    """The chunk digest applied to an array of groups: [g, k] uint64 or [g, k] uint8 -> [g]."""
    a = np.ascontiguousarray(words)
    a = a.reshape(len(a), -1).view(np.uint8)
    pad = (-a.shape[1]) % 8
    if pad:
        a = np.concatenate([a, np.zeros((len(a), pad), np.uint8)], 1)
    w = a.view('<u8')
    with np.errstate(over='ignore'):
        def mix(z):
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            return z ^ (z >> np.uint64(31))
        i = np.arange(1, w.shape[1] + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        h = mix(w + i[None, :]).sum(axis=1, dtype=np.uint64)
        return mix(h ^ np.uint64(a.shape[1]))


def served_requests():
    """[(rate, seed, chunks), ...]: the checked prefix of each fixed-rate stream."""
    out = []
    pool = workloads.seed_pool()
    for rate in workloads.SERVE_RATES:
        reqs = workloads.requests(rate, 10.0 * workloads.SERVE_CHECKED / rate, pool=pool)[:workloads.SERVE_CHECKED]
        assert len(reqs) == workloads.SERVE_CHECKED
        out += [(rate, s, ch) for _, s, ch in reqs]
    return out


def c3(R, threads):
    t0 = time.time()
    work = workloads.batched()
    seeds = np.array([s for s, ch in work for _ in ch], np.int64)
    xz = np.array([c for _, ch in work for c in ch], np.int32)
    dt, db = R.digest_terrain(seeds, xz, threads=threads)
    np.savez_compressed(os.path.join(HERE, 'c3_digests.npz'), seed=seeds, xz=xz, terrain=dt, biomes=db)
    print('c3: %d chunks, %.0f s' % (len(xz), time.time() - t0), flush=True)


def c4(R, threads):
    t0 = time.time()
    reqs = served_requests()
    seeds = np.array([s for _, s, ch in reqs for _ in ch], np.int64)
    xz = np.array([c for _, _, ch in reqs for c in ch], np.int32)
    counts = np.array([len(ch) for _, _, ch in reqs], np.int32)
    rates = np.array([r for r, _, _ in reqs], np.float64)
    dt, db = R.digest_terrain(seeds, xz, threads=threads)
    np.savez_compressed(os.path.join(HERE, 'c4_digests.npz'), rate=rates, request_chunks=counts, seed=seeds, xz=xz, terrain=dt, biomes=db)
    print('c4: %d requests, %d chunks, %.0f s' % (len(reqs), len(xz), time.time() - t0), flush=True)


def edge(R, threads):
    t0 = time.time()
    folds = workloads.fold_chunks()
    border = workloads.border_region()
    seeds = np.array([workloads.SEED] * len(folds) + [1] * len(folds) + [workloads.SEED] * len(border), np.int64)
    xz = np.array(folds + folds + border, np.int32)
    dt, db = R.digest_terrain(seeds, xz, threads=threads)
    np.savez_compressed(os.path.join(HERE, 'edge_digests.npz'), seed=seeds, xz=xz, terrain=dt, biomes=db)
    print('edge: %d chunks, %.0f s' % (len(xz), time.time() - t0), flush=True)


def c5_queries(R, threads, state):
    path = os.path.join(state, 'c5_queries_%d.npy' % QUERIES)
    if os.path.exists(path):
        return np.load(path)
    t0 = time.time()
    seeds, _, _ = workloads.stress()
    qs, q = workloads.queries(QUERIES, seeds)
    order = np.argsort(qs, kind='stable')        # by seed: each thread keeps its seed's tables
    ans = np.zeros(QUERIES, np.uint8)
    ans[order] = R.query_biomes(qs[order], q[order], threads=threads, run=256)
    np.save(path, ans)
    print('c5 queries: %d, %.0f s' % (QUERIES, time.time() - t0), flush=True)
    return ans


def c5_chunks(R, threads, state, slice_seeds=32):
    seeds, which, xz = workloads.stress()
    n = len(xz)
    order = np.argsort(which, kind='stable')
    dt, db = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
    bounds = np.searchsorted(which[order], np.arange(0, len(seeds) + slice_seeds, slice_seeds))
    t0 = time.time()
    for k in range(len(bounds) - 1):
        lo, hi = int(bounds[k]), int(bounds[k + 1])
        if lo >= hi:
            continue
        idx = order[lo:hi]
        path = os.path.join(state, 'c5_%04d.npz' % k)
        if os.path.exists(path):
            f = np.load(path)
            assert np.array_equal(f['index'], idx)
            dt[idx], db[idx] = f['terrain'], f['biomes']
            continue
        a, b = R.digest_terrain(seeds[which[idx]], xz[idx], threads=threads, run=8)
        np.savez(path, index=idx, terrain=a, biomes=b)
        dt[idx], db[idx] = a, b
        print('c5 chunks: slice %d / %d (%d chunks), %.0f s' % (k + 1, len(bounds) - 1, hi, time.time() - t0), flush=True)
    return dt, db


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--threads', type=int, default=os.cpu_count())
    ap.add_argument('--only', default='c3,c4,c5q,c5,edge')
    ap.add_argument('--state', default='/tmp/mct_digest_state')
    a = ap.parse_args()
    parts = set(a.only.split(','))
    os.makedirs(a.state, exist_ok=True)
    R = oracle.Reference()
    if 'c3' in parts:
        c3(R, a.threads)
    if 'c4' in parts:
        c4(R, a.threads)
    if 'edge' in parts:
        edge(R, a.threads)
    if 'c5' in parts:
        dt, db = c5_chunks(R, a.threads, a.state)
    ans = c5_queries(R, a.threads, a.state) if ('c5q' in parts or 'c5' in parts) else None
    if 'c5' in parts:
        np.savez_compressed(os.path.join(HERE, 'c5_digests.npz'), group=np.int64(GROUP), qgroup=np.int64(QGROUP), queries=np.int64(QUERIES),
                            terrain_groups=fold(dt.reshape(-1, GROUP)), biome_groups=fold(db.reshape(-1, GROUP)),
                            query_groups=fold(ans.reshape(-1, QGROUP)),
                            head_terrain=dt[:GROUP], head_biomes=db[:GROUP], head_answers=ans[:QGROUP])
        print('c5 written', flush=True)


if __name__ == '__main__':
    main()
