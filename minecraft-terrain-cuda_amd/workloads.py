"""The chunk lists of the measured and tested workloads, generated deterministically from
SplitMix64 (constants below), so the bench, the tests and the fixture writer see the same
inputs everywhere.

  plain_region_set   one seed, three 64 x 64-chunk regions (origin, (-2048, -2048), and one
                     ending at chunk 117,187 = the chunk of block 1,874,999) + the noise-wrap
                     chunks + the reference's world-border chunks (chunk +-1,874,999)
  c1_region          8 x 8 chunks at (0, 0) under one seed drawn from SplitMix64
  batched            256 uniform 64-bit seeds, each with a log-uniform 1..1024 chunk list,
                     half contiguous near-square rectangles, half scattered in +-117,187
  requests           served stream: Poisson arrivals, seed from a 64-seed pool, W x H region
                     (W, H uniform 1..32) in spiral order around a random player chunk
  stress             1,024 seeds, 2^20 chunks uniform over +-117,187 (wrap chunks included)
  queries            biome-only block queries: seed from the stress seeds, x / z uniform over
                     the blocks of +-117,187 chunks, y uniform in -64..319
  fold_chunks        one chunk on each side of the noise lattice's wrap fold for every distinct
                     horizontal scale x frequency product of the IR's noise sampling
  border_region      the reference's 8 x 8 border verify region at chunk (1874992, 1874992)
"""
import math
import os

MASK = (1 << 64) - 1
SEED = 123456789              # the reference bench's seed (bench/bench_main.cpp:50)
BORDER_CHUNK = 1874999 >> 4   # 117,187: chunk of block 1,874,999
CHUNK_RANGE = BORDER_CHUNK
# chunks straddling the noise lattice's 32-bit wrap (GradientNoise.wrap folds), and the
# reference's own world-border chunks (chunk coordinate +-1,874,999)
WRAP_CHUNKS = [(6128, 0), (-6129, 5), (11, 6128), (-3, -6129)]
REF_BORDER_CHUNKS = [(1874999, 1874999), (-1874999, 1874998), (1874999, -1874999), (-1874999, -1874999)]
EDGE_CASES = [(123456789, [(0, 0)]), (0, [(-2048, -2048)]), (1, [(6128, 0), (-6129, 5)]),
              (-1, [(1874999, -1874999), (11, 6128)]), (-4172144997902289642, [(-3, -6129), (7, -9)])]

C1_STREAM = 0x43315F5345454453     # SplitMix64 start for the C1 seed
BATCH_STREAM = 0x4241544348454421  # 256-seed batch
SERVE_STREAM = 0x5345525645212121  # request stream
STRESS_STREAM = 0x5354524553532121
# The served configuration's three fixed Poisson rates (requests/s): 25 / 50 / 80% of the
# plain generator's request capacity on this mix (about 400 requests/s, 108,808 chunks/s,
# when they were fixed). Each rate's stream is prefix-stable, so the checked requests are the
# first ones the bench serves.
SERVE_RATES = (100.0, 200.0, 320.0)
SERVE_CHECKED = 40    # requests per rate whose every chunk is checked against reference digests


class SplitMix64:
    def __init__(self, state):
        self.s = state & MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def signed(self):
        v = self.next()
        return v - (1 << 64) if v >> 63 else v

    def uniform(self):            # [0, 1)
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def below(self, n):           # [0, n)
        return (self.next() >> 11) % n

    def between(self, lo, hi):    # [lo, hi]
        return lo + self.below(hi - lo + 1)


BORDER_VERIFY = (1874992, 8)   # bench/run_all.py:27: x0 = z0 = 1874992, size 8
FOLD = float(1 << 24)          # GradientNoise.wrap folds a lattice coordinate at |d| = 2^24


def fold_products(ir_path=None):
    """Distinct horizontal products (block coordinate -> lattice coordinate) of the IR's noise
    sampling: xz_scale x octave frequency for every noise node (both NormalNoise forks), 1/4 x
    frequency for the offset noise of shift_a / shift_b, and BlendedNoise's 684.412 x xz_scale
    x octave factor for its limit noises and / xz_factor for its main noise."""
    import json
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, 'codegen'))
    import model
    with open(ir_path or os.path.join(here, 'data', 'synthetic_overworld_ir.json')) as fh:
        ir = json.load(fh)
    out = set()
    for n in ir['nodes']:
        if n['t'] in ('noise', 'shift_a', 'shift_b'):
            s = model.f64h(n['xz_scale']) if n['t'] == 'noise' else 0.25
            for _, fr, _ in model.normal_noise_layers(ir['noises'][n['noise']]):
                out.update((s * fr, s * fr * model.SECOND_FREQ))
        elif n['t'] == 'blended_noise':
            xz = 684.412 * model.f64h(n['xz_scale'])
            for first, div in ((-15, 1.0), (-7, model.f64h(n['xz_factor']))):
                out.update(xz / div * f for f, _, _ in model.fbm_layers(first, 0.0, 1.0))
    return sorted(out)


def fold_chunks(limit=1874999):
    """For each fold product p whose fold lies within chunk +-limit: the chunk holding block
    2^24 / p on the +x axis and the chunk holding block -2^24 / p on the -z axis (the other
    coordinate a small offset, so the chunks are distinct)."""
    out = []
    for i, p in enumerate(fold_products()):
        b = FOLD / p
        c = math.floor(b / 16.0)
        if c > limit:
            continue
        out.append((c, i % 7))
        out.append((i % 5 - 2, math.floor(-b / 16.0)))
    return out


def border_region():
    x0, w = BORDER_VERIFY
    return region(x0, x0, w)


def region(x0, z0, w, d=None):
    d = w if d is None else d
    return [(x, z) for z in range(z0, z0 + d) for x in range(x0, x0 + w)]


def plain_region_set(side=64):
    chunks = region(0, 0, side) + region(-2048, -2048, side) + region(BORDER_CHUNK - side + 1, -BORDER_CHUNK, side)
    return chunks + WRAP_CHUNKS + REF_BORDER_CHUNKS


def c1_seed():
    return SplitMix64(C1_STREAM).signed()


def c1_region():
    return c1_seed(), region(0, 0, 8)


def batched(n_seeds=256, max_chunks=1024):
    """[(seed, [(cx, cz), ...]), ...]: even entries are near-square rectangles (the first n
    chunks of a ceil(sqrt n)-wide rectangle), odd entries n scattered chunks."""
    r = SplitMix64(BATCH_STREAM)
    out = []
    for i in range(n_seeds):
        seed = r.signed()
        n = int(round(math.exp(r.uniform() * math.log(max_chunks))))
        n = max(1, min(max_chunks, n))
        if i % 2 == 0:
            w = int(math.ceil(math.sqrt(n)))
            h = (n + w - 1) // w
            x0 = r.between(-CHUNK_RANGE, CHUNK_RANGE - w + 1)
            z0 = r.between(-CHUNK_RANGE, CHUNK_RANGE - h + 1)
            chunks = region(x0, z0, w, h)[:n]
        else:
            chunks = [(r.between(-CHUNK_RANGE, CHUNK_RANGE), r.between(-CHUNK_RANGE, CHUNK_RANGE)) for _ in range(n)]
        out.append((seed, chunks))
    return out


def spiral(cx, cz, w, h):
    """The W x H region around (cx, cz) (the player chunk at its centre), nearest rings first;
    within a ring in the order of the angle, starting east."""
    x0, z0 = cx - (w - 1) // 2, cz - (h - 1) // 2
    pts = region(x0, z0, w, h)
    return sorted(pts, key=lambda p: (max(abs(p[0] - cx), abs(p[1] - cz)), math.atan2(p[1] - cz, p[0] - cx) % (2 * math.pi), p))


def seed_pool(n=64):
    r = SplitMix64(SERVE_STREAM ^ 0x504F4F4C)
    return [r.signed() for _ in range(n)]


def requests(rate_per_s, duration_s, pool=None, max_side=32):
    """Open-loop Poisson stream: [(arrival_s, seed, [(cx, cz), ...]), ...]."""
    pool = seed_pool() if pool is None else pool
    r = SplitMix64(SERVE_STREAM + int(rate_per_s * 1000))
    t, out = 0.0, []
    while True:
        t += -math.log(1.0 - r.uniform()) / rate_per_s
        if t >= duration_s:
            return out
        seed = pool[r.below(len(pool))]
        w, h = r.between(1, max_side), r.between(1, max_side)
        px, pz = r.between(-CHUNK_RANGE + max_side, CHUNK_RANGE - max_side), r.between(-CHUNK_RANGE + max_side, CHUNK_RANGE - max_side)
        out.append((t, seed, spiral(px, pz, w, h)))


def stress_seeds(n_seeds=1024):
    r = SplitMix64(STRESS_STREAM)
    return [r.signed() for _ in range(n_seeds)]


def splitmix_array(start, n):
    """Outputs 1..n of SplitMix64 from `start`, vectorised (numpy uint64 wraps mod 2^64)."""
    import numpy as np
    with np.errstate(over='ignore'):
        z = np.uint64(start & MASK) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def stress(n_seeds=1024, n_chunks=1 << 20):
    """(seeds, chunk_seed_index, chunks) as numpy arrays: uniform chunks over +-117,187 with
    the wrap chunks at the front, seed index uniform over the seed list."""
    import numpy as np
    seeds = np.array(stress_seeds(n_seeds), np.int64)
    v = splitmix_array(STRESS_STREAM ^ 0x43484E4B, n_chunks)
    span = np.uint64(2 * CHUNK_RANGE + 1)
    xz = np.empty((n_chunks, 2), np.int32)
    xz[:, 0] = ((v & np.uint64(0xFFFFFF)) % span).astype(np.int64) - CHUNK_RANGE
    xz[:, 1] = (((v >> np.uint64(24)) & np.uint64(0xFFFFFF)) % span).astype(np.int64) - CHUNK_RANGE
    which = ((v >> np.uint64(48)) % np.uint64(n_seeds)).astype(np.int32)
    xz[:len(WRAP_CHUNKS)] = WRAP_CHUNKS
    which[:len(WRAP_CHUNKS)] = 0
    return seeds, which, xz


def queries(n, seeds):
    """n biome-only block queries: (seed per query, xyz int32 [n, 3])."""
    import numpy as np
    v = splitmix_array(STRESS_STREAM ^ 0x51554552, 2 * n).reshape(n, 2)
    lo, span = -CHUNK_RANGE * 16, np.uint64((2 * CHUNK_RANGE + 1) * 16)
    out = np.empty((n, 3), np.int32)
    out[:, 0] = ((v[:, 0] & np.uint64(0xFFFFFFFF)) % span).astype(np.int64) + lo
    out[:, 1] = ((v[:, 1] & np.uint64(0xFFFF)) % np.uint64(384)).astype(np.int64) - 64
    out[:, 2] = ((v[:, 0] >> np.uint64(32)) % span).astype(np.int64) + lo
    s = np.asarray(seeds, np.int64)[((v[:, 1] >> np.uint64(16)) % np.uint64(len(seeds))).astype(np.int64)]
    return s, out
