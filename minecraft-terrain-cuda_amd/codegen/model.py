"""Seed-independent model data derived from the IR and terrain model at build time.

Everything here is a pure function of the two JSON files, so it is computed once by the code
generator and compiled into the library:

* the noise layer plan: per NormalNoise key its kept octaves, two layers each (the second at
  frequency x 1.0181268882175227), with the normalised value factor; per BlendedNoise node
  the 16 + 16 + 8 smeared fbm layers (seed_tables.cpp:23-184 in the reference states the
  construction);
* the recipe each layer's octave is drawn with (which random stream, which fork, which name),
  and the MD5 digests of all names (RandomSupport.seedFromHashOf);
* the climate R-tree (Climate.RTree build: 19 children per node, stable sorts), flattened
  with every node's children stored contiguously in DFS-visit order (biome.cpp:95-199);
* the three legacy-random simplex permutations of biome temperature (terrain.cpp:70-80);
* palette flags, biome climate, special indices.
"""
import hashlib
import json
import math
import struct

SECOND_FREQ = 1.0181268882175227


def f32(x):
    return struct.unpack('<f', struct.pack('<f', float(x)))[0]


def f32b(bits):
    return struct.unpack('<f', struct.pack('<I', bits & 0xFFFFFFFF))[0]


def f64h(h):
    return struct.unpack('<d', struct.pack('<Q', int(h, 16)))[0]


def md5_longs(name):
    d = hashlib.md5(name.encode('utf-8')).digest()
    return int.from_bytes(d[:8], 'big'), int.from_bytes(d[8:], 'big')


# ---- NormalNoise / BlendedNoise layer plan -------------------------------------------------
def java_stream_sum(values):
    """DoubleStream.sum(): compensated (Kahan-style) summation."""
    s0 = s1 = simple = 0.0
    for v in values:
        tmp = v - s1
        velvel = s0 + tmp
        s1 = (velvel - s0) - tmp
        s0 = velvel
        simple += v
    r = s0 - s1
    if math.isnan(r) and math.isinf(simple):
        return simple
    return r


def normal_noise_layers(d):
    """NormalNoise(params) construction order: returns [(octave_index, frequency, value_factor)]
    for each kept octave; each becomes two layers."""
    base_octave = d['base_octave']
    n = d['octave_count']
    mods = [f64h(m) for m in d['amplitude_modifiers']]
    mod = (lambda i: mods[i]) if mods else (lambda i: 1.0)
    normalize = d['normalize'] != 'disabled'
    base_amp = f64h(d['base_amplitude'])
    freq = math.ldexp(1.0, base_octave)
    amp = base_amp
    if normalize:
        amp *= math.ldexp(1.0, n - 1) / (math.ldexp(1.0, n) - 1.0)
    octs = []
    for i in range(n):
        m = mod(i)
        if m != 0.0:
            octs.append((base_octave + i, freq, amp * m))
        freq *= 2.0
        amp *= 0.5
    variance = 0.0
    for _, _, a in octs:
        layer = 0.2702247831245211 * abs(a)
        variance += layer * layer
    dev = math.sqrt(variance)
    if dev == 0.0:
        norm = 0.0
    else:
        target = java_stream_sum([abs(a) for _, _, a in octs])
        norm = (target * 0.3333333333333333) / (dev * math.sqrt(2.0))
    if d['normalize'] == 'legacy' and norm != 0.0:
        lo = min((i for i in range(n) if mod(i) != 0.0), default=2147483647)
        hi = max((i for i in range(n) if mod(i) != 0.0), default=-2147483648)
        span = hi - lo
        expected = 0.1 * (1.0 + 1.0 / (span + 1))
        norm = base_amp * 0.5 * 0.3333333333333333 / expected
    return [(idx, fr, f32(norm * a)) for idx, fr, a in octs]


def fbm_layers(first_octave, smear_y, value_factor):
    """BlendedNoise.createFbm: layers from the coarsest factor 1.0 downwards."""
    octaves = -first_octave + 1
    factor = 1.0
    value_factor /= math.ldexp(1.0, octaves) - 1.0
    out = []
    for _ in range(octaves):
        out.append((factor, f32(value_factor), smear_y * factor))
        factor /= 2.0
        value_factor *= 2.0
    return out


class LayerPlan:
    """Global layer order: NormalNoise keys sorted, each key's layers in construction order;
    then each BlendedNoise node's min_limit, max_limit and main sets."""

    def __init__(self, ir):
        self.keys = sorted(ir['noises'].keys())
        self.key_index = {k: i for i, k in enumerate(self.keys)}
        self.freq, self.amp, self.smear = [], [], []
        self.recipe = []          # (kind, a, b, c): kind 0 normal (key, octave_index, fork); 1 blended (node_k, position)
        self.noise_range = []     # (first, count) per key
        self.octave_names = []
        for k in self.keys:
            first = len(self.freq)
            for idx, fr, vf in normal_noise_layers(ir['noises'][k]):
                name = 'octave_%d' % idx
                if name not in self.octave_names:
                    self.octave_names.append(name)
                for fork, fmul in ((0, 1.0), (1, SECOND_FREQ)):
                    self.freq.append(fr * fmul if fork else fr)
                    self.amp.append(vf)
                    self.smear.append(0.0)
                    self.recipe.append((0, self.key_index[k], self.octave_names.index(name), fork))
            self.noise_range.append((first, len(self.freq) - first))
        self.blended_ids = sorted(n['id'] for n in ir['nodes'] if n['t'] == 'blended_noise')
        self.blended = []         # per node: dict set -> (first, count)
        for bk, nid in enumerate(self.blended_ids):
            n = ir['nodes'][nid]
            y_mult = 684.412 * f64h(n['y_scale'])
            limit_smear = y_mult * f64h(n['smear_scale_multiplier'])
            main_smear = limit_smear / f64h(n['y_factor'])
            limit_factor = f32(0.99998474)
            sets = {}
            pos = 0
            for name, fo, sm, vf in (('min_limit', -15, limit_smear, limit_factor), ('max_limit', -15, limit_smear, limit_factor),
                                     ('main', -7, main_smear, 12.75)):
                first = len(self.freq)
                for fr, a, s in fbm_layers(fo, sm, vf):
                    self.freq.append(fr)
                    self.amp.append(a)
                    self.smear.append(s)
                    self.recipe.append((1, bk, pos, 0))
                    pos += 1
                sets[name] = (first, len(self.freq) - first)
            self.blended.append(sets)
        self.count = len(self.freq)


# ---- climate R-tree -----------------------------------------------------------------------
def _center(a, b):
    s = a + b
    return s // 2 if s >= 0 else -((-s) // 2)


class RTree:
    """Climate.RTree build (stable sorts; buckets of 19^floor(log(n - .01) / log 19))."""

    def __init__(self, params):
        self.mins, self.maxs, self.kids, self.leaf = [], [], [], []
        leaves = []
        for i, p in enumerate(params):
            self.mins.append(list(p[0]))
            self.maxs.append(list(p[1]))
            self.kids.append([])
            self.leaf.append(i)
            leaves.append(i)
        self.root = self._build(leaves)

    def _subtree(self, kids):
        mn = [min(self.mins[k][d] for k in kids) for d in range(7)]
        mx = [max(self.maxs[k][d] for k in kids) for d in range(7)]
        self.mins.append(mn)
        self.maxs.append(mx)
        self.kids.append(list(kids))
        self.leaf.append(-1)
        return len(self.mins) - 1

    def _c(self, n, d):
        return _center(self.mins[n][d], self.maxs[n][d])

    def _build(self, children):
        if not children:
            raise ValueError('empty R-tree node')
        if len(children) == 1:
            return children[0]
        if len(children) <= 19:
            children = sorted(children, key=lambda n: sum(abs(self._c(n, d)) for d in range(7)))
            return self._subtree(children)
        lg = math.floor(math.log(float(len(children)) - 0.01) / math.log(19.0))
        expected = int(math.pow(19.0, lg))
        min_cost, min_dim, min_buckets = None, -1, None
        work = list(children)
        for d in range(7):
            work.sort(key=lambda n: tuple(self._c(n, (d + k) % 7) for k in range(7)))
            buckets = [work[i:i + expected] for i in range(0, len(work), expected)]
            total = 0
            for b in buckets:
                for dd in range(7):
                    total += abs(max(self.maxs[k][dd] for k in b) - min(self.mins[k][dd] for k in b))
            if min_cost is None or min_cost > total:
                min_cost, min_dim, min_buckets = total, d, [list(b) for b in buckets]
        bucket_nodes = [self._subtree(b) for b in min_buckets]
        bucket_nodes.sort(key=lambda n: tuple(abs(self._c(n, (min_dim + k) % 7)) for k in range(7)))
        built = [self._build(self.kids[bn]) for bn in bucket_nodes]
        return self._subtree(built)

    def flatten(self):
        """Nodes renumbered so each node's children are contiguous, root first; returns
        (mins, maxs, first_child, child_count, leaf)."""
        order = [self.root]
        first = {}
        i = 0
        while i < len(order):
            n = order[i]
            if self.kids[n]:
                first[n] = len(order)
                order.extend(self.kids[n])
            i += 1
        new = {n: k for k, n in enumerate(order)}
        mins = [self.mins[n] for n in order]
        maxs = [self.maxs[n] for n in order]
        fc = [new[self.kids[n][0]] if self.kids[n] else -1 for n in order]
        cc = [len(self.kids[n]) for n in order]
        lf = [self.leaf[n] for n in order]
        return mins, maxs, fc, cc, lf


def quantize(v):
    """Climate.quantizeCoord((float) v): (long)(v * 10000.0F)."""
    return int(f32(f32(v) * f32(10000.0)))


# ---- legacy random (java.util.Random) for the temperature simplex permutations -------------
class LegacyRandom:
    def __init__(self, seed):
        self.s = (seed ^ 0x5DEECE66D) & ((1 << 48) - 1)

    def next(self, bits):
        self.s = (self.s * 0x5DEECE66D + 0xB) & ((1 << 48) - 1)
        v = self.s >> (48 - bits)           # (int) of a value below 2^32
        return v - (1 << 32) if v >= 1 << 31 else v

    def next_int(self, bound):
        if bound & (bound - 1) == 0:
            return (bound * self.next(31)) >> 31
        while True:
            sample = self.next(31)
            modulo = sample % bound
            if ((sample - modulo + (bound - 1)) & 0xFFFFFFFF) < 0x80000000:
                return modulo

    def next_double(self):
        return ((self.next(26) << 27) + self.next(27)) * 2.0 ** -53


def simplex_perm(r):
    """SimplexNoise(random): three offset draws (unused by the temperature noise) and the
    256-entry permutation."""
    for _ in range(3):
        r.next_double()
    p = list(range(256))
    for i in range(256):
        o = r.next_int(256 - i)
        p[i], p[o + i] = p[o + i], p[i]
    return p


def temperature_perms():
    r1 = LegacyRandom(1234)
    t = simplex_perm(r1)
    r3 = LegacyRandom(3456)
    frozen = [simplex_perm(r3) for _ in range(3)]
    r2 = LegacyRandom(2345)
    info = simplex_perm(r2)
    return t, frozen, info


def load_params(tm):
    out = []
    for p in tm['biome_parameters']:
        mn = [p['p'][2 * d] for d in range(6)] + [p['offset']]
        mx = [p['p'][2 * d + 1] for d in range(6)] + [p['offset']]
        out.append((mn, mx, p['biome']))
    return out


def file_hash(path):
    """FNV-1a 64 of the file bytes (identifies the IR / terrain model the library was built from)."""
    h = 0xcbf29ce484222325
    with open(path, 'rb') as fh:
        data = fh.read()
    for b in data:
        h ^= b
        h = (h * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h
