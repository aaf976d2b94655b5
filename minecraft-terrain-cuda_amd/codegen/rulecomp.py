"""Compiler of the terrain model's material-rule tree (MaterialSystem / SurfaceRules) into
one inlined function, `apply_rules(c)`, over a surface context `c`.

Vanilla's rules return a block state or nothing: a sequence returns its first rule's
non-empty result, a condition rule its follow-up's result when the condition holds
(`cpu/src/terrain.cpp:454-746`, Surface::apply). The compiler turns that into structured
fall-through code: each rule is a block of statements that either returns a state or falls
through to whatever follows it. A sequence is its rules' blocks one after another, a
condition an `if` around its follow-up's block, a block state `return s;`, an ore vein a
return when the vein yields a state. Code after a block that always returns is dropped.

Every condition kind is one method of the surface context (the pipeline's Surf), so this
module emits only the tree's structure and constants. While compiling, it also collects
facts the pipeline uses, each conservative by construction (a path condition can only narrow):

  ore y range     per ore-vein rule, the hull over its paths of the y intervals the plain
                  y_above conditions on the path allow (kOreYLo / kOreYHi)
  column noises   2D noise thresholds read the column record; per slot, the union over the
                  paths to its conditions of the biome sets the paths require
                  (kColNoiseBiomes), so a column none of whose candidate biomes is in it
                  leaves the slot unread
"""
from densitygraph import f64_from_hex
from lowering import dlit

ALL_BIOMES = (1 << 128) - 1
SURFACE_SLOTS = ('minecraft:surface', 'minecraft:surface_secondary', 'minecraft:clay_bands_offset')


def hex64(v):
    return '0x%016Xull' % (v & 0xFFFFFFFFFFFFFFFF)


def deref(c):
    while c['t'] == 'ref':
        c = c['cond']
    return c


class Facts:
    __slots__ = ('ylo', 'yhi', 'biomes')

    def __init__(self, ylo, yhi, biomes=ALL_BIOMES):
        self.ylo, self.yhi, self.biomes = ylo, yhi, biomes

    def under(self, cond):
        """The facts inside a condition rule's follow-up."""
        c = deref(cond)
        f = Facts(self.ylo, self.yhi, self.biomes)
        if plain_y(c):
            f.ylo = max(f.ylo, c['anchor'])
        elif c['t'] == 'biome':
            f.biomes &= biome_mask(c)
        elif c['t'] == 'not':
            inner = deref(c['cond'])
            if plain_y(inner):
                f.yhi = min(f.yhi, inner['anchor'])
            elif inner['t'] == 'biome':
                f.biomes &= ~biome_mask(inner) & ALL_BIOMES
        return f


def plain_y(c):
    return c['t'] == 'y_above' and not c['add_stone_depth'] and c['surface_depth_multiplier'] == 0


def biome_mask(c):
    m = 0
    for b in c['biomes']:
        if not 0 <= b < 128:
            raise SystemExit('biome index %d outside 0..127' % b)
        m |= 1 << b
    return m


class RuleCompiler:
    def __init__(self, plan, layers):
        self.plan, self.P = plan, layers
        g = plan.g
        self.ymin, self.yend = g.min_y, g.min_y + g.height
        self.slots = list(SURFACE_SLOTS)
        self.slot_biomes = {k: ALL_BIOMES for k in range(len(self.slots))}
        self.random_names = []
        self.ore_y = {}

    def stack(self, key):
        if key not in self.P.key_index:
            raise SystemExit('rule noise %s is not in the IR noise list' % key)
        return self.P.noise_range[self.P.key_index[key]]

    def slot(self, key):
        self.stack(key)
        if key not in self.slots:
            self.slots.append(key)
        return self.slots.index(key)

    def cond(self, c, facts):
        c = deref(c)
        t = c['t']
        stone = 'true' if c.get('add_stone_depth') else 'false'
        if t == 'biome':
            m = biome_mask(c)
            return 'c.biome_in(%s, %s)' % (hex64(m), hex64(m >> 64))
        if t == 'noise_threshold':
            lo, hi = dlit(f64_from_hex(c['min'])), dlit(f64_from_hex(c['max']))
            if not c['is_3d']:
                k = self.slot(c['noise'])
                self.slot_biomes[k] = self.slot_biomes.get(k, 0) | facts.biomes
                return 'c.noise_in(c.col_noise(%d), %s, %s)' % (k, lo, hi)
            first, count = self.stack(c['noise'])
            return 'c.noise_in(mct::stack_point<%d, %d>(c.O, static_cast<double>(c.x), static_cast<double>(c.y), static_cast<double>(c.z)), %s, %s)' % (
                first, count, lo, hi)
        if t == 'vertical_gradient':
            if c['random_name'] not in self.random_names:
                self.random_names.append(c['random_name'])
            return 'c.vertical_gradient(%d, %d, %d)' % (self.random_names.index(c['random_name']), c['true_at_and_below'], c['false_at_and_above'])
        if t == 'y_above':
            return 'c.y_above(%d, %d, %s)' % (c['anchor'], c['surface_depth_multiplier'], stone)
        if t == 'water':
            return 'c.water_above(%d, %d, %s)' % (c['offset'], c['surface_depth_multiplier'], stone)
        if t == 'not':
            return '!(%s)' % self.cond(c['cond'], facts)
        if t == 'stone_depth':
            return 'c.stone_depth(%d, %s, %d, %s)' % (c['offset'], 'true' if c['add_surface_depth'] else 'false', c['secondary_depth_range'],
                                                       'true' if c['ceiling'] else 'false')
        if t == 'temperature':
            return 'c.cold_enough_to_snow()'
        if t == 'steep':
            return 'c.steep()'
        if t == 'hole':
            return 'c.hole()'
        if t == 'above_preliminary_surface':
            return 'c.above_preliminary_surface()'
        raise SystemExit('unsupported surface condition %r' % t)

    def rule(self, r, facts, out, ind):
        """Appends rule r's statements; True when they always return."""
        t = r['t']
        if t == 'ref':
            return self.rule(r['rule'], facts, out, ind)
        if t == 'block':
            out.append('%sreturn %d;' % (ind, r['state']))
            return True
        if t == 'bandlands':
            out.append('%sreturn c.band();' % ind)
            return True
        if t == 'ore_vein':
            k = self.plan.ore_rules.index(r)
            if facts.ylo < facts.yhi:
                lo, hi = self.ore_y.get(k, (facts.ylo, facts.yhi))
                self.ore_y[k] = (min(lo, facts.ylo), max(hi, facts.yhi))
            out.append('%s{ const int s = c.ore_vein(%d); if (s >= 0) return s; }' % (ind, k))
            return False
        if t == 'condition':
            out.append('%sif (%s) {' % (ind, self.cond(r['if'], facts)))
            self.rule(r['then'], facts.under(r['if']), out, ind + '    ')
            out.append('%s}' % ind)
            return False
        if t == 'sequence':
            for x in r['rules']:
                if self.rule(x, facts, out, ind):
                    return True
            return False
        raise SystemExit('unsupported surface rule %r' % t)

    def compile(self):
        body = []
        ends = self.rule(self.plan.terrain['material_rule'], Facts(self.ymin, self.yend), body, '    ')
        o = ['template <class C> MCT_HD MCT_FI int apply_rules(C& c) {'] + body
        if not ends:
            o.append('    return -1;')
        o.append('}')
        nore = max(len(self.plan.ore_rules), 1)
        rng = [self.ore_y.get(k, (0, 0)) for k in range(nore)]
        o.append('// per ore-vein rule: the y range [lo, hi) its blocks can be reached at (empty: never)')
        o.append('constexpr int kOreYLo[%d] = {%s};' % (nore, ', '.join(str(a) for a, _ in rng)))
        o.append('constexpr int kOreYHi[%d] = {%s};' % (nore, ', '.join(str(b) for _, b in rng)))
        o.append('constexpr int kColNoises = %d;   // 2D noises of the column record' % len(self.slots))
        o.append('constexpr int kColSurface = 0, kColSecondary = 1, kColClay = 2;')
        o.append('// per column-noise slot: the biomes under which the rules can read it (bit b of word b >> 6)')
        o.append('constexpr uint64_t kColNoiseBiomes[kColNoises][2] = {%s};' % ', '.join(
            '{%s, %s}' % (hex64(self.slot_biomes.get(k, 0)), hex64(self.slot_biomes.get(k, 0) >> 64)) for k in range(len(self.slots))))
        o.append('MCT_HD MCT_FI float surf_column_noise(const mct::Octave* O, double x, double z, int k) {')
        o.append('    switch (k) {')
        for k, key in enumerate(self.slots):
            first, count = self.stack(key)
            o.append('    case %d: return mct::stack_point<%d, %d>(O, x, 0.0, z);' % (k, first, count))
        o.append('    }')
        o.append('    return 0.0f;')
        o.append('}')
        return o

