"""Throughput of the MI355X TERRAIN-step generator.

A step is one pass of the full TERRAIN step (biomes, density + aquifer fill, material rules
with ore veins) over the plain single-seed workload: three 64 x 64-chunk regions (origin,
chunk (-2048, -2048), and one ending at chunk 117,187, the chunk of block 1,874,999) plus the
noise-wrap chunks and the reference's world-border chunks, with outputs kept in HBM. `value`
is chunks/s of that step on one GPU. Alongside it the line carries:

  e2e_host_chunks_per_s   the same chunks with results copied to pinned host memory
  batched                 mixed-seed batch (256 uniform seeds, uneven chunk lists), chunks/s
  served                  Poisson request streams: p50 / p99 latency and chunks/s per rate
  stress                  1,024 seeds x 2^20 scattered chunks, chunks/s; 2^24 biome queries/s
  roofline                the dominant kernel's VALU issue rate against the MI355X VALU peak
  cpu_baseline            the reference cpu/ port (built from its sources) on the host cores
  stages_ms / kernels_ms  per-stage and per-kernel times of one step (HIP events)

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--quick]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'minecraft-terrain-cuda_amd'))

from workloads import SEED, BORDER_CHUNK, region, plain_region_set  # noqa: E402

# MI355X VALU peak: the published FP32 vector peak, 157.3 TFLOP/s, is 256 CUs x 128 lanes x
# 2.4 GHz x 2 flops (FMA). In instructions: 32-bit VALU operations issue at two wave64
# instructions per CU per cycle (128 lane-ops), FP64 ones at half that (the published FP64
# vector peak, 78.6 TFLOP/s). The roofline counts VALU lane-ops against the 32-bit rate; a
# kernel with FP64 work cannot reach 1.0.
VALU_PEAK_TOPS = 256 * 128 * 2.4e9 / 1e12

# Work per unit of the main kernels, measured with rocprofv3 on this workload (the summaries
# are in profiles/valu_plain3x64.md and profiles/hbm_plain3x64.md; per thread there, per unit
# here):
#   VALU lane-ops   SQ_INSTS_VALU x 64 (the instruction stream the kernel issues)
#   HBM bytes       FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the factors from a plain 1 GiB
#                   device-to-device copy measured in the same runs
WORK_PER_UNIT = {
    'surface2': {'unit': 'block column', 'valu': 12505.6, 'hbm': 360.31 + 76.36},
    # per corner evaluated: the launch covers every corner column slot of the batch, 1.62
    # threads per distinct corner on this workload, and the idle workgroups exit at once
    'corners': {'unit': 'corner', 'valu': 17988.6, 'hbm': 31.29},
    # 2 restricted group searches per quart column + per quart its final minimum and target
    'biome_quarts': {'unit': 'quart column of a biome chunk', 'valu': 2 * 23538.4 + 96 * (69.2 + 150.0),
                     'hbm': 2 * (31.95 + 16.00) + 96 * (58.05 + 12.00 + 1.51 + 56.05 + 4.00)},
    'fill': {'unit': 'block column of a cell (8 blocks)', 'valu': 1141.7, 'hbm': 11.29 + 13.15},
    'ore_veins': {'unit': 'block column', 'valu': 11352.4, 'hbm': 125.28 + 259.67},
}


def distinct_columns(chunks):
    cols = set()
    for x, z in chunks:
        for k in range(25):
            cols.add((x * 4 + k % 5, z * 4 + k // 5))
    return len(cols)


def units_per_step(chunks, batch):
    """Threads (units) each main kernel runs over the whole workload, batch by batch (corner
    columns and biome chunks are deduplicated within a batch)."""
    u = {'surface2': 256 * len(chunks), 'fill': 256 * 48 * len(chunks), 'ore_veins': 256 * len(chunks), 'corners': 0, 'biome_quarts': 0}
    for k in range(0, len(chunks), batch):
        part = chunks[k:k + batch]
        u['corners'] += distinct_columns(part) * 49
        u['biome_quarts'] += len({(x + dx, z + dz) for x, z in part for dz in (-1, 0, 1) for dx in (-1, 0, 1)}) * 16
    return u


def cpu_baseline(seconds_target=15.0):
    """The reference cpu/ port (oracle/_ref, built from its sources) in its bench's region
    mode (biomes of the region + border once, then fill + surface per chunk) on a thread
    pool of the host's cores."""
    lib = os.path.join(ROOT, 'oracle', '_ref', 'libmcref.so')
    if not os.path.exists(lib):
        return None
    import mct
    P, I64, I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    R = ctypes.CDLL(lib)
    R.ref_create.restype = P
    R.ref_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    R.ref_generate_region.argtypes = [P, I64, I, I, I, I, I, P, P]
    h = R.ref_create(mct.IR_PATH.encode(), mct.TERRAIN_PATH.encode())
    threads = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else os.cpu_count()
    threads = max(1, min(threads, 16))
    w = 8
    t = np.zeros(w * w * 98304, np.uint16)
    t0 = time.perf_counter()
    R.ref_generate_region(h, SEED, 0, 0, w, w, threads, None, t.ctypes.data)
    dt = time.perf_counter() - t0
    # grow the region to about seconds_target of work
    w2 = int(max(8, min(64, w * (seconds_target / max(dt, 1e-3)) ** 0.5)))
    t = np.zeros(w2 * w2 * 98304, np.uint16)
    t0 = time.perf_counter()
    R.ref_generate_region(h, SEED, 0, 0, w2, w2, threads, None, t.ctypes.data)
    dt = time.perf_counter() - t0
    return {'value': round(w2 * w2 / dt, 2), 'unit': 'chunks/s', 'cores': threads, 'kind': 'reference',
            'sample': '%dx%d-chunk region at (0,0), seed %d, reference cpu/ port region mode, %.1f s' % (w2, w2, SEED, dt)}


def batched_vs_plain(mct, batch, tuning=None):
    """The 256-seed uneven workload in one mixed-seed call against the same chunks generated
    plainly, one seed's list per call (outputs in HBM both ways)."""
    import workloads
    work = workloads.batched()
    seeds = np.array([s for s, ch in work for _ in ch], np.int64)
    chunks = [c for _, ch in work for c in ch]
    n = len(chunks)
    d_t, d_b = mct.alloc_device(n * mct.BLOCKS_PER_CHUNK), mct.alloc_device(n * mct.QUARTS_PER_CHUNK)
    g = mct.FullTerrainGenerator(max_batch_chunks=batch, pipeline_slots=2, seed_capacity=256)
    if tuning is not None:
        g.autotune(seeds[:batch], chunks[:batch])        # shapes for this mix, before the timed calls
    g.generate_device(seeds, chunks, d_t, d_b)          # tables resident, buffers sized
    t0 = time.perf_counter()
    g.generate_device(seeds, chunks, d_t, d_b)
    tb = time.perf_counter() - t0
    t0 = time.perf_counter()
    off = 0
    for s, ch in work:                                  # plain: one seed per call
        g.generate_device(np.full(len(ch), s, np.int64), ch, d_t + off * mct.BLOCKS_PER_CHUNK, d_b + off * mct.QUARTS_PER_CHUNK)
        off += len(ch)
    tp = time.perf_counter() - t0
    g.close()
    mct.free_device(d_t)
    mct.free_device(d_b)
    return {'workload': '256 uniform seeds, log-uniform 1..1024 chunks each (half near-square rectangles, half scattered in +-117,187)',
            'chunks': n, 'seeds': len(work), 'chunks_per_s': round(n / tb, 1), 'plain_chunks_per_s': round(n / tp, 1), 'speedup': round(tp / tb, 2)}


def served(mct, plain_rate, seconds=4.0, fractions=(0.25, 0.5, 0.8, 3.0), max_batch=1024):
    """Open-loop Poisson request streams at fractions of the plain generator's capacity on the
    same request mix: p50 / p99 latency (submit to results in host memory) and throughput. The
    last rate (3x) overloads the server on purpose: its chunks/s is the served capacity."""
    import threading
    import workloads
    pool = workloads.seed_pool()
    mean_chunks = 16.5 * 16.5
    ring_chunks = 40000
    ring = mct.PinnedBuffer((ring_chunks, 384, 16, 16))
    bring = mct.PinnedBuffer((ring_chunks, 96, 4, 4))
    out = []
    srv = mct.Server(max_batch_chunks=max_batch, seed_capacity=64, hbm_budget_bytes=48 << 30)
    for f in fractions:
        rps = f * plain_rate / mean_chunks
        reqs = workloads.requests(rps, seconds, pool=pool)
        srv.poll()
        pos = [0]
        sub = {}

        def client():
            t0 = time.perf_counter()
            for t, seed, chunks in reqs:
                dt = t - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                k = len(chunks)
                if pos[0] + k > ring_chunks:
                    pos[0] = 0                           # the ring is far larger than what is in flight
                sub[srv.submit(seed, chunks, ring.array[pos[0]:pos[0] + k], bring.array[pos[0]:pos[0] + k])] = k
                pos[0] += k
        t0 = time.perf_counter()
        th = threading.Thread(target=client)
        th.start()
        th.join()
        srv.drain()
        wall = time.perf_counter() - t0
        lat = np.array([l for _, l in srv.poll(1 << 20)])
        nch = sum(len(c) for _, _, c in reqs)
        out.append({'offered_fraction': f, 'offered_requests_per_s': round(rps, 1), 'requests': len(reqs), 'chunks': nch,
                    'chunks_per_s': round(nch / wall, 1), 'p50_ms': round(float(np.percentile(lat, 50)), 3),
                    'p99_ms': round(float(np.percentile(lat, 99)), 3), 'max_ms': round(float(lat.max()), 3)})
    st = srv.stats()
    srv.close()
    ring.free()
    bring.free()
    return {'workload': 'Poisson arrivals, seed from a 64-seed pool, W x H region (W, H uniform 1..32) in spiral order around a random player chunk',
            'plain_capacity_chunks_per_s': round(plain_rate, 1), 'seconds_per_rate': seconds, 'rates': out,
            'served_capacity_chunks_per_s': out[-1]['chunks_per_s'], 'served_vs_plain': round(out[-1]['chunks_per_s'] / plain_rate, 2),
            'batches': st['batches'], 'mixed_batches': st['mixed_batches'], 'max_batch': max_batch}


def stress(mct, batch, window=1 << 16, n_queries=1 << 24):
    """The mixed stress workload: 2^20 chunks uniform over +-117,187 (wrap chunks included) of
    1,024 seeds, streamed through a window of HBM output buffers (the whole output, ~105 GB,
    is not kept); then 2^24 biome-only block queries (climate sampling, R-tree search chain
    and zoom per query) from host arrays, answers back in host memory."""
    import workloads
    seeds, which, xz = workloads.stress()
    cs = seeds[which]
    n = len(xz)
    d_t, d_b = mct.alloc_device(window * mct.BLOCKS_PER_CHUNK), mct.alloc_device(window * mct.QUARTS_PER_CHUNK)
    g = mct.FullTerrainGenerator(max_batch_chunks=batch, pipeline_slots=2, seed_capacity=1024)
    g.generate_device(cs[:window], xz[:window], d_t, d_b)   # every seed's tables resident, buffers sized
    t0 = time.perf_counter()
    for off in range(0, n, window):
        g.generate_device(cs[off:off + window], xz[off:off + window], d_t, d_b)
    tc = time.perf_counter() - t0
    mct.free_device(d_t)
    mct.free_device(d_b)
    qs, q = workloads.queries(n_queries, seeds)
    g.biome_query(qs[:1 << 16], q[:1 << 16])
    t0 = time.perf_counter()
    g.biome_query(qs, q)
    tq = time.perf_counter() - t0
    g.close()
    return {'workload': '1,024 seeds, 2^20 chunks uniform over +-117,187 (wrap chunks first), %d-chunk HBM output window; '
                        '2^24 biome-only block queries (x, z uniform over the same blocks, y uniform -64..319)' % window,
            'chunks': n, 'chunks_per_s': round(n / tc, 1), 'seconds': round(tc, 3),
            'queries': n_queries, 'queries_per_s': round(n_queries / tq, 1), 'query_seconds': round(tq, 3)}


def plain_request_capacity(mct, n_requests=200):
    """Chunks/s of the plain generator on the served request mix: one request per call,
    results copied to pinned host memory (what a server without batching does)."""
    import workloads
    pool = workloads.seed_pool()
    reqs = workloads.requests(1000.0, 10.0, pool=pool)[:n_requests]
    g = mct.FullTerrainGenerator(max_batch_chunks=1024, pipeline_slots=2, seed_capacity=64)
    buf = mct.PinnedBuffer((1024, 384, 16, 16))
    for s in pool:
        g.set_seed(s)
        g.generate([(0, 0)], out=buf.array[:1])
    t0 = time.perf_counter()
    nch = 0
    for _, seed, chunks in reqs:
        g.set_seed(seed)
        g.generate(chunks, out=buf.array[:len(chunks)])
        nch += len(chunks)
    dt = time.perf_counter() - t0
    g.close()
    buf.free()
    return nch / dt


def copy_calibration(nbytes):
    """One plain device-to-device copy of nbytes (HIP's blit kernel): under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE against nbytes give the scale of those counters."""
    import mct
    hip = ctypes.CDLL('libamdhip64.so')
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    src, dst = mct.alloc_device(nbytes), mct.alloc_device(nbytes)
    try:
        rc = hip.hipMemcpy(dst, src, nbytes, 3)   # hipMemcpyDeviceToDevice
        rc = rc or hip.hipDeviceSynchronize()
        if rc != 0:
            raise RuntimeError('hipMemcpy device to device failed: %d' % rc)
    finally:
        mct.free_device(src)
        mct.free_device(dst)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--batch', type=int, default=0, help='chunks per device batch (0: the fastest of 4096 / 8192 / 16384 on this workload)')
    ap.add_argument('--slots', type=int, default=2)
    ap.add_argument('--quick', action='store_true', help='one 32x32 region instead of the full workload')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-serve', action='store_true', help='skip the served-request measurement')
    ap.add_argument('--no-stress', action='store_true', help='skip the 1,024-seed stress leg')
    ap.add_argument('--no-tune', action='store_true', help='default launch shapes (no autotuning)')
    ap.add_argument('--tuning-json', metavar='JSON', help='launch shapes to use instead of autotuning, as the "tuning" object of a bench line')
    ap.add_argument('--plain-only', action='store_true', help='only the plain workload (no batched or served legs), e.g. under a profiler')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='BYTES',
                    help='also run one device-to-device copy of BYTES (the reference point for HBM counters)')
    a = ap.parse_args()
    if a.gpus != 1:
        print('bench.py measures one GPU: this workload shards by chunk tiles with no exchange step; '
              'runs on more than one GPU are not part of this build (--gpus %d refused)' % a.gpus, file=sys.stderr)
        sys.exit(2)
    import mct
    chunks = region(0, 0, 32) if a.quick else plain_region_set()
    n = len(chunks)
    seeds = np.full(n, SEED, np.int64)
    d_t = mct.alloc_device(n * mct.BLOCKS_PER_CHUNK)
    d_b = mct.alloc_device(n * mct.QUARTS_PER_CHUNK)
    batch_trials = None
    if a.batch == 0:   # batch size: one untimed pass per candidate, before any timed step
        batch_trials = {}
        for b in (4096, 8192, 16384):
            gt = mct.FullTerrainGenerator(max_batch_chunks=b, pipeline_slots=a.slots, seed_capacity=256)
            gt.generate_device(seeds, chunks, d_t, d_b)
            t0 = time.perf_counter()
            gt.generate_device(seeds, chunks, d_t, d_b)
            batch_trials[b] = round(n / (time.perf_counter() - t0), 1)
            gt.close()
        a.batch = max(batch_trials, key=batch_trials.get)
    g = mct.FullTerrainGenerator(max_batch_chunks=a.batch, pipeline_slots=a.slots, seed_capacity=256)
    tuning = None
    if a.tuning_json:
        tuning = json.loads(a.tuning_json)
        g.set_tuning(tuning)
    elif not a.no_tune:   # launch shapes timed on the workload's first batch, before any timed step
        tuning = g.autotune(seeds[:a.batch], chunks[:a.batch])
    for _ in range(a.warmup):
        g.generate_device(seeds, chunks, d_t, d_b)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        g.generate_device(seeds, chunks, d_t, d_b)
    dt = (time.perf_counter() - t0) / a.steps
    value = n / dt
    # one profiled step on a single-slot generator (no overlap between batches, so each
    # kernel's HIP-event time is its own): stage and kernel times
    gp = mct.FullTerrainGenerator(max_batch_chunks=a.batch, pipeline_slots=1, seed_capacity=256)
    if tuning:
        gp.set_tuning(tuning)
    gp.generate_device(seeds, chunks, d_t, d_b)
    tm = mct.Timing()
    gp.profile(True)
    gp.generate_device(seeds, chunks, d_t, d_b, timing=tm)
    kt = gp.kernel_times()
    gp.profile(False)
    gp.close()
    stages = {k: round(v, 3) for k, v in tm.as_dict().items()}
    kernels = {k: round(v[0], 3) for k, v in kt.items() if v[1] > 0}
    dom = max(kernels, key=kernels.get)
    # roofline of the dominant kernel: its work per launch over its average launch time (HIP
    # events on the stream it runs on, single-slot pass above)
    launches = {k: v[1] for k, v in kt.items()}
    units = units_per_step(chunks, a.batch)
    w = WORK_PER_UNIT.get(dom)
    roof = {'bound': 'valu', 'kernel': dom, 'achieved': None, 'peak': VALU_PEAK_TOPS, 'unit': 'TOP/s', 'frac': None, 'traffic': None}
    if w and dom in units:
        per_launch = units[dom] / launches[dom]
        avg_s = kernels[dom] * 1e-3 / launches[dom]
        achieved = w['valu'] * per_launch / avg_s / 1e12
        roof.update({'achieved': round(achieved, 3), 'frac': round(achieved / VALU_PEAK_TOPS, 4), 'traffic': round(w['hbm'] * per_launch),
                     'avg_launch_ms': round(avg_s * 1e3, 4), 'units_per_launch': round(per_launch, 1), 'work_unit': w['unit'],
                     'valu_per_unit': w['valu'], 'hbm_gbs': round(w['hbm'] * per_launch / avg_s / 1e9, 2)})
    # end to end with results copied to pinned host memory
    pin = mct.lib().mct_alloc_pinned(n * mct.BLOCKS_PER_CHUNK)
    buf = np.ctypeslib.as_array(ctypes.cast(pin, ctypes.POINTER(ctypes.c_uint8)), shape=(n, 384, 16, 16))
    g.set_seed(SEED)
    g.generate(chunks, biomes=False, out=buf)
    t0 = time.perf_counter()
    g.generate(chunks, biomes=False, out=buf)
    e2e = n / (time.perf_counter() - t0)
    mct.lib().mct_free_pinned(pin)
    mct.free_device(d_t)
    mct.free_device(d_b)
    g.close()
    batched = None if (a.quick or a.plain_only) else batched_vs_plain(mct, a.batch, None if a.no_tune else tuning)
    serve = None
    if not a.quick and not a.no_serve and not a.plain_only:
        serve = served(mct, plain_request_capacity(mct))
    mixed = None if (a.quick or a.plain_only or a.no_stress) else stress(mct, a.batch)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
    cpu = None if a.no_cpu else cpu_baseline()
    line = {'metric': 'chunks/s full TERRAIN, batched/served vs plain, 1 MI355X; p50/p99 latency', 'value': round(value, 1), 'unit': 'chunks/s',
            'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': round(dt * 1e3, 3), 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': 'f32+f64+i64', 'data': 'synthetic (deterministic synthetic overworld IR and terrain model)',
            'config': {'workload': 'plain TERRAIN, one seed (%d), %s, full height, outputs in HBM' % (
                SEED, '32x32 region' if a.quick else '3 x 64x64-chunk regions (origin, (-2048,-2048), border chunk %d) + 8 wrap/border chunks' % BORDER_CHUNK),
                'chunks': n, 'batch': a.batch, 'slots': a.slots},
            'e2e_host_chunks_per_s': round(e2e, 1), 'batch_trials': batch_trials, 'tuning': tuning, 'batched': batched, 'served': serve, 'stress': mixed, 'roofline': roof, 'cpu_baseline': cpu,
            'stages_ms': stages, 'kernels_ms': kernels}
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
