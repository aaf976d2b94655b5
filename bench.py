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
  roofline                the dominant kernel's counted algorithmic operations (FP64 / FP32 /
                          INT32, from the code generator's per-scope counts) per launch over its
                          HIP-event launch time, against MI355X's published vector peaks
  cpu_baseline            the reference cpu/ port (built from its sources) on every host core
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

# MI355X published vector peaks (TFLOP/s; an FMA counts 2): FP64 78.6, FP32 157.3. The FP32
# figure is 128 FMA lanes per CU per cycle (256 CUs, 2.4 GHz): one VALU lane-op per lane slot,
# 78.6 T lane-ops/s, is the issue ceiling the counted operations are set against. Only a few
# instructions (32-bit add, and, move, FP32 multiply) issue at that rate on the MI355X; most
# (min / max, shifts, 24-bit multiply-adds, med3, compares, selects, conversions, FP64) issue
# at half of it, one wave instruction per SIMD every 4 cycles. This path's arithmetic is
# unfused (Java semantics: no contraction), so an FP32 operation is one lane-op and one flop,
# at most half of what the FP32 peak counts.
PEAK_FP64 = 78.6
PEAK_FP32 = 157.3
LANE_OPS_PER_S = PEAK_FP32 / 2 * 1e12

# device kernel group -> the emulation stage that counts its operations (one to one)
KERNEL_STAGE = {'fill': 'fill', 'corners': 'corners', 'columns': 'columns', 'fts': 'fts', 'aquifer_cells': 'aquifer_cells',
                'biome_columns': 'biome_columns', 'biome_quarts': 'biome_quarts', 'surface_columns': 'surface_columns'}


def count_ops(seed, chunks, threads, group=256):
    """Algorithmic operations per stage over the whole workload, counted by the counting host
    emulation in groups of `group` consecutive chunks on a thread pool:
    {stage: {'unit': u, 'per_unit': {class: ops per unit}}}."""
    import gen_host
    from concurrent.futures import ThreadPoolExecutor
    parts = [chunks[k:k + group] for k in range(0, len(chunks), group)]
    with ThreadPoolExecutor(max_workers=threads) as ex:
        results = list(ex.map(lambda part: gen_host.op_counts(seed, part), parts))
    tot = {}
    for r in results:
        for st, v in r.items():
            t = tot.setdefault(st, {'unit': v['unit'], 'units': 0, 'ops': dict.fromkeys(gen_host.OP_CLASSES, 0)})
            t['units'] += v['units']
            for c in gen_host.OP_CLASSES:
                t['ops'][c] += v['ops'][c]
    return {st: {'unit': t['unit'], 'per_unit': {c: t['ops'][c] / max(t['units'], 1) for c in gen_host.OP_CLASSES}} for st, t in tot.items()}


def batch_units(seeds, chunks, batch):
    """Units (chunks, distinct biome chunks, corner columns, grid keys) summed over the
    workload's device batches."""
    import gen_host
    out = {}
    for k in range(0, len(chunks), batch):
        for u, v in gen_host.stage_units(seeds[k:k + batch], chunks[k:k + batch]).items():
            out[u] = out.get(u, 0) + v
    return out


def roofline(kernels, launches, dom, counts, units, n, value):
    """The dominant kernel's counted operations per launch over its average HIP-event launch
    time, against the published peaks: frac = FP64 rate / 78.6 T + FP32 rate / 157.3 T, the
    share of the launch the counted floating-point work needs at peak; `achieved` is the same
    as FP32-equivalent TFLOP/s (FP64 weighted by the peak ratio) so that frac = achieved /
    peak. issue_frac: all counted operations (FP64 + FP32 + INT32) against the VALU lane-op
    ceiling. `path`: the whole step's counted operations per chunk at `value`."""
    roof = {'bound': 'valu', 'kernel': dom, 'achieved': None, 'peak': PEAK_FP32, 'unit': 'TFLOP/s', 'frac': None, 'traffic': None}
    stage = KERNEL_STAGE.get(dom)
    if stage is None or stage not in counts or not launches.get(dom):
        roof['note'] = 'no operation count for kernel group %s' % dom
        return roof
    avg_s = kernels[dom] * 1e-3 / launches[dom]
    per = counts[stage]['per_unit']
    unit = counts[stage]['unit']
    per_launch = units[unit] / launches[dom]
    ops = {c: per[c] * per_launch for c in per}
    f64, f32, i32 = ops['f64'] / avg_s / 1e12, ops['f32'] / avg_s / 1e12, ops['i32'] / avg_s / 1e12
    achieved = f32 + f64 * PEAK_FP32 / PEAK_FP64
    roof.update({
        'achieved': round(achieved, 3), 'frac': round(achieved / PEAK_FP32, 4),
        'avg_launch_ms': round(avg_s * 1e3, 4), 'launches': launches[dom], 'work_unit': unit, 'units_per_launch': round(per_launch, 1),
        'ops_per_unit': {c: round(v, 1) for c, v in per.items()},
        'fp64': {'achieved': round(f64, 3), 'peak': PEAK_FP64, 'frac': round(f64 / PEAK_FP64, 4)},
        'fp32': {'achieved': round(f32, 3), 'peak': PEAK_FP32, 'frac': round(f32 / PEAK_FP32, 4)},
        'int32_tops': round(i32, 3),
        'issue_frac': round((ops['f64'] + ops['f32'] + ops['i32']) / avg_s / LANE_OPS_PER_S, 4)})
    tot = {c: sum(v['per_unit'][c] * units[v['unit']] for v in counts.values()) / n for c in ('f64', 'f32', 'i32', 'mem')}
    p64, p32 = tot['f64'] * value / 1e12, tot['f32'] * value / 1e12
    roof['path'] = {'ops_per_chunk': {c: round(v) for c, v in tot.items()}, 'fp64_tflops': round(p64, 3), 'fp32_tflops': round(p32, 3),
                    'frac': round(p64 / PEAK_FP64 + p32 / PEAK_FP32, 4),
                    'uncounted': 'surface material-rule walks, dedupe sorts, seed-table setup'}
    return roof


def traffic(dom, chunks_per_launch):
    """HBM bytes per launch of kernel group `dom` from the committed counter profile of this
    workload (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, scaled by the same counters over a
    plain device-to-device copy), per chunk x chunks per launch; None without one."""
    path = os.path.join(ROOT, 'profiles', 'hbm_%s_plain3x64.json' % dom)
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        p = json.load(fh)
    return round(p['bytes_per_chunk'] * chunks_per_launch), os.path.relpath(path, ROOT)


def note(msg):
    """One progress line on stderr as each leg ends (the JSON line stays alone on stdout)."""
    print('bench: %s' % msg, file=sys.stderr, flush=True)


def host_cores():
    """Logical cores this process may run on (its affinity set), the thread count of the CPU
    baseline."""
    return len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else (os.cpu_count() or 1)


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
    threads = host_cores()
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
    """The 256-seed uneven workload in one mixed-seed call, with every seed's tables built
    inside the call (cold) and then with them resident, against the same chunks generated
    plainly, one seed's list per call (outputs in HBM every way)."""
    import workloads
    work = workloads.batched()
    seeds = np.array([s for s, ch in work for _ in ch], np.int64)
    chunks = [c for _, ch in work for c in ch]
    n = len(chunks)
    d_t, d_b = mct.alloc_device(n * mct.BLOCKS_PER_CHUNK), mct.alloc_device(n * mct.QUARTS_PER_CHUNK)
    g = mct.FullTerrainGenerator(max_batch_chunks=batch, pipeline_slots=2, seed_capacity=256)
    if tuning is not None:
        g.autotune(seeds[:batch], chunks[:batch])        # shapes for this mix, before the timed calls
    # buffers sized and batch shapes captured on the same chunk lists under 256 other seeds,
    # which fill the seed cache: the first timed call builds every table itself
    g.generate_device(seeds ^ np.int64(0x5DEECE66D), chunks, d_t, d_b)
    t0 = time.perf_counter()
    g.generate_device(seeds, chunks, d_t, d_b)
    tc = time.perf_counter() - t0
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
            'chunks': n, 'seeds': len(work), 'chunks_per_s': round(n / tb, 1), 'cold_chunks_per_s': round(n / tc, 1),
            'plain_chunks_per_s': round(n / tp, 1), 'speedup': round(tp / tb, 2), 'speedup_cold': round(tp / tc, 2)}


def served(mct, plain_rate, seconds=4.0, overload=3.0, max_batch=1024):
    """Open-loop Poisson request streams at the served configuration's three fixed rates
    (requests/s) and one overload rate (`overload` x the plain generator's capacity on the same
    mix): p50 / p99 latency (submit to results in host memory) and throughput. The overload
    stream keeps the server saturated: its chunks/s is the served capacity."""
    import threading
    import workloads
    pool = workloads.seed_pool()
    mean_chunks = 16.5 * 16.5
    ring_chunks = 40000
    ring = mct.PinnedBuffer((ring_chunks, 384, 16, 16))
    bring = mct.PinnedBuffer((ring_chunks, 96, 4, 4))
    out = []
    srv = mct.Server(max_batch_chunks=max_batch, seed_capacity=64, hbm_budget_bytes=48 << 30)
    for rps in list(workloads.SERVE_RATES) + [overload * plain_rate / mean_chunks]:
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
        out.append({'offered_fraction': round(rps * mean_chunks / plain_rate, 3), 'offered_requests_per_s': round(rps, 1), 'requests': len(reqs), 'chunks': nch,
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
    # buffers sized and shapes captured under 1,024 other seeds, so that the timed pass builds
    # every table itself (setup included)
    g.generate_device(cs[:window] ^ np.int64(0x5DEECE66D), xz[:window], d_t, d_b)
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
    return {'workload': '1,024 seeds (tables built in the timed pass), 2^20 chunks uniform over +-117,187 (wrap chunks first), %d-chunk HBM output window; '
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
        note('batch trials %s' % batch_trials)
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
    note('plain: %.1f chunks/s' % value)
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
    # roofline of the dominant kernel: its counted operations per launch over its average
    # launch time (HIP events on the stream it runs on, single-lane pass above)
    launches = {k: v[1] for k, v in kt.items()}
    units = batch_units(seeds, chunks, a.batch)
    counts = count_ops(SEED, chunks, host_cores())
    roof = roofline(kernels, launches, dom, counts, units, n, value)
    roof['traffic'], src = traffic(dom, units['chunks'] / max(launches.get(dom, 1), 1))
    note('profiled pass and operation counts done')
    if src:
        roof['traffic_source'] = src
    # end to end with results copied to pinned host memory
    pin = mct.lib().mct_alloc_pinned(n * mct.BLOCKS_PER_CHUNK)
    buf = np.ctypeslib.as_array(ctypes.cast(pin, ctypes.POINTER(ctypes.c_uint8)), shape=(n, 384, 16, 16))
    g.set_seed(SEED)
    g.generate(chunks, biomes=False, out=buf)
    t0 = time.perf_counter()
    g.generate(chunks, biomes=False, out=buf)
    e2e = n / (time.perf_counter() - t0)
    # the host link's ceiling for the same bytes: one device-to-pinned-host copy of the terrain
    mct._check(mct.lib().mct_device_to_host(pin, d_t, n * mct.BLOCKS_PER_CHUNK))
    t0 = time.perf_counter()
    mct._check(mct.lib().mct_device_to_host(pin, d_t, n * mct.BLOCKS_PER_CHUNK))
    d2h_s = time.perf_counter() - t0
    d2h = {'gb_per_s': round(n * mct.BLOCKS_PER_CHUNK / d2h_s / 1e9, 2), 'chunks_per_s': round(n / d2h_s, 1)}
    mct.lib().mct_free_pinned(pin)
    note('host output: %.1f chunks/s' % e2e)
    mct.free_device(d_t)
    mct.free_device(d_b)
    g.close()
    batched = None if (a.quick or a.plain_only) else batched_vs_plain(mct, a.batch, None if a.no_tune else tuning)
    if batched:
        note('batched: %s' % batched['chunks_per_s'])
    serve = None
    if not a.quick and not a.no_serve and not a.plain_only:
        serve = served(mct, plain_request_capacity(mct))
        note('served done')
    mixed = None if (a.quick or a.plain_only or a.no_stress) else stress(mct, a.batch)
    if mixed:
        note('stress done')
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
    cpu = None if a.no_cpu else cpu_baseline()
    line = {'metric': 'chunks/s full TERRAIN, batched/served vs plain, 1 MI355X; p50/p99 latency', 'value': round(value, 1), 'unit': 'chunks/s',
            'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': round(dt * 1e3, 3), 'higher_is_better': True, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': 'f32+f64+i64', 'data': 'synthetic (deterministic synthetic overworld IR and terrain model)',
            'config': {'workload': 'plain TERRAIN, one seed (%d), %s, full height, outputs in HBM' % (
                SEED, '32x32 region' if a.quick else '3 x 64x64-chunk regions (origin, (-2048,-2048), border chunk %d) + 8 wrap/border chunks' % BORDER_CHUNK),
                'chunks': n, 'batch': a.batch, 'slots': a.slots},
            'e2e_host_chunks_per_s': round(e2e, 1), 'd2h_ceiling': d2h, 'batch_trials': batch_trials, 'tuning': tuning, 'batched': batched, 'served': serve, 'stress': mixed, 'roofline': roof, 'cpu_baseline': cpu,
            'stages_ms': stages, 'kernels_ms': kernels}
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
