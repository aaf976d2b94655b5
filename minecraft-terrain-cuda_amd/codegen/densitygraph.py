"""The density-function graph of a `mccuda-ir/1` file, and the pipeline plan derived from it.

The IR is a dense post-order list of nodes (children before parents; `cpu/include/mccuda/ir.h:
12-83`, `cpu/src/ir.cpp:63-157`). This module gives the code generator three things:

  DensityGraph   typed access: operands of a node in a fixed order, literal decoding
                 (f32 as u32 bits, f64 as hex bits), and a check of the graph's shape
  PipelinePlan   which nodes each pipeline stage evaluates: the interpolation channels of the
                 fill (4 x 8 cells), the y = 0 column slots of the corner and climate stages,
                 the point-semantics roots (aquifer noises, ore filler gaps), the single
                 find_top_surface node, and the ore-vein rules of the terrain model
  binary_decision  vanilla's compile-time choice for a binary node (constant operand order,
                 MIN/MAX pruning from value ranges, division by a constant as a multiply):
                 the decisions `cpu/src/volume_eval.cpp:38-71` states

The plan enforces the structural constraints the pipeline relies on (one find_top_surface
node feeding `aquifer/surface_level` as a y = 0 slice and `surface/chunk_surface_level` as a
slice of a 16 x 1 interpolation; all channels 4 x 8).
"""
import struct

F32_BITS = 0xFFFFFFFF


def f32_from_bits(bits):
    return struct.unpack('<f', struct.pack('<I', bits & F32_BITS))[0]


def f64_from_hex(h):
    return struct.unpack('<d', struct.pack('<Q', int(h, 16)))[0]


class GraphError(SystemExit):
    pass


class DensityGraph:
    # operand fields of each node kind, in evaluation order (the order only fixes the text;
    # every node is a pure function of position, so any order gives the same values)
    OPERANDS = {
        'cache': ('in',), 'blend_density': ('in',), 'slice': ('in',), 'interpolated': ('in',), 'unary': ('in',), 'clamp': ('in',),
        'binary': ('l', 'r'), 'lerp': ('alpha', 'first', 'second'), 'range_choice': ('in', 'when_in', 'when_out'),
        'find_top_surface': ('density', 'upper_bound'),
    }

    def __init__(self, doc):
        if doc.get('format') != 'mccuda-ir/1':
            raise GraphError('unexpected IR format %r' % doc.get('format'))
        self.doc = doc
        self.nodes = doc['nodes']
        self.roots = doc['roots']
        self.min_y, self.height, self.sea_level = doc['min_y'], doc['height'], doc['sea_level']
        for i, n in enumerate(self.nodes):
            if n['id'] != i:
                raise GraphError('IR node ids must be dense and in order (node %d)' % i)
            for c in self.operands(i):
                if not 0 <= c < i:
                    raise GraphError('node %d reads node %d, which does not precede it' % (i, c))

    def kind(self, i):
        return self.nodes[i]['t']

    def root(self, name):
        if name not in self.roots:
            raise GraphError('the IR has no root %r' % name)
        return self.roots[name]

    def operands(self, i):
        """Distinct operand node ids of node i (first occurrence order)."""
        n = self.nodes[i]
        t = n['t']
        out = [n[f] for f in self.OPERANDS.get(t, ())]
        if t == 'noise':
            out += [n['shift_' + a] for a in 'xyz' if not n['shift_%s_zero' % a]]
        elif t == 'interval_select':
            out = [n['in']] + list(n['functions'])
        elif t == 'spline':
            out = spline_inputs(n['spline'])
        seen, uniq = set(), []
        for c in out:
            if c not in seen:
                seen.add(c)
                uniq.append(c)
        return uniq

    def value_range(self, i):
        r = self.nodes[i]['range']
        return f32_from_bits(r[0]), f32_from_bits(r[1])

    def is_const(self, i):
        return self.kind(i) == 'constant'

    def const_value(self, i):
        return f32_from_bits(self.nodes[i]['v'])

    def is_y0_slice(self, i):
        n = self.nodes[i]
        return n['t'] == 'slice' and n['axis'] == 'y' and n['coord'] == 0

    def frontier(self, roots, kind, through=lambda i: True):
        """Nodes of `kind` reachable from roots without passing through one (depth first,
        operand order, first visit order); `through(i)` False stops the walk at node i."""
        out, seen = [], set()

        def visit(i):
            if i in seen:
                return
            seen.add(i)
            if self.kind(i) == kind:
                out.append(i)
                return
            if not through(i):
                return
            for c in self.operands(i):
                visit(c)
        for r in roots:
            visit(r)
        return out


def spline_inputs(s):
    """Coordinate nodes a (possibly nested) CubicSpline reads, first occurrence order."""
    out = []

    def walk(p):
        if 'coord' in p:
            if p['coord'] not in out:
                out.append(p['coord'])
            for v in p['val']:
                walk(v)
    walk(s)
    return out


def binary_decision(g, i):
    """(op, a, b) for binary node i: op in add / sub / mul / div / divc (divide by the constant
    b) / min / max / pass (a alone decides: its range is entirely below / above the other's).
    For min / max with a constant left operand the operands are swapped, so a constant is
    always the second operand (vanilla's TwoArgumentSimpleFunction normalisation)."""
    n = g.nodes[i]
    l, r, op = n['l'], n['r'], n['op']
    if op in ('add', 'sub', 'mul'):
        return op, l, r
    if op == 'div':
        return ('divc' if g.is_const(r) and not g.is_const(l) else 'div'), l, r
    if op not in ('min', 'max'):
        raise GraphError('unknown binary op %r (node %d)' % (op, i))
    (lmin, lmax), (rmin, rmax) = g.value_range(l), g.value_range(r)
    if op == 'min':
        if lmax < rmin:
            return 'pass', l, None
        if rmax < lmin:
            return 'pass', r, None
    else:
        if lmin > rmax:
            return 'pass', l, None
        if rmin > lmax:
            return 'pass', r, None
    return (op, r, l) if g.is_const(l) else (op, l, r)


CLIMATE = ('temperature', 'vegetation', 'continents', 'erosion', 'depth', 'ridges')
AQUIFER_POINT = ('barrier', 'fluid_level_floodedness', 'fluid_level_spread', 'lava', 'exclusion')


class PipelinePlan:
    def __init__(self, g, terrain):
        if terrain.get('format') != 'mccuda-terrain/1':
            raise GraphError('unexpected terrain model format %r' % terrain.get('format'))
        if (terrain['sea_level'], terrain['gen_min_y'], terrain['gen_depth']) != (g.sea_level, g.min_y, g.height):
            raise GraphError('IR and terrain model disagree on the world height / sea level')
        self.g = g
        self.terrain = terrain
        self.final = g.root('final_density')
        self.ore_rules = ore_vein_rules(terrain['material_rule'])
        self.ore_roots = []
        for o in self.ore_rules:
            self.ore_roots += [g.root(o['density']), g.root(o['richness'])]
        self.climate = [g.root('climate/' + k) for k in CLIMATE]
        tops = [i for i in range(len(g.nodes)) if g.kind(i) == 'find_top_surface']
        if len(tops) != 1:
            raise GraphError('the IR must have exactly one find_top_surface node (it has %d)' % len(tops))
        self.fts = tops[0]
        fts = g.nodes[self.fts]
        self.fts_density, self.fts_upper = fts['density'], fts['upper_bound']
        self.fts_lower, self.fts_cell = fts['lower_bound'], fts['cell_height']
        a = g.root('aquifer/surface_level')
        if not (g.is_y0_slice(a) and g.nodes[a]['in'] == self.fts):
            raise GraphError('aquifer/surface_level must be a y = 0 slice of find_top_surface')
        c = g.nodes[g.root('surface/chunk_surface_level')]
        inner = g.nodes[c['in']] if c['t'] == 'slice' else None
        if not (inner and inner['t'] == 'interpolated' and inner['in'] == self.fts and (inner['cell_xz'], inner['cell_y']) == (16, 1)):
            raise GraphError('surface/chunk_surface_level must be a slice of a 16 x 1 interpolation of find_top_surface')
        self.point_roots = [g.root('aquifer/' + k) for k in AQUIFER_POINT] + [g.root(o['filler_gap']) for o in self.ore_rules]
        # interpolation channels: the fill's (final density), then the ore evaluators' others
        self.channels = g.frontier([self.final], 'interpolated')
        self.ore_channels = [x for x in g.frontier(self.ore_roots, 'interpolated') if x not in self.channels]
        for x in self.channels + self.ore_channels:
            n = g.nodes[x]
            if (n['cell_xz'], n['cell_y']) != (4, 8):
                raise GraphError('interpolation channel %d is not a 4 x 8 cell' % x)
        self.all_channels = self.channels + self.ore_channels
        # column slots: y = 0 slices the corner stage reads below its channels, and those of
        # the climate roots; both are evaluated once per 4-block column
        corner_inputs = [g.nodes[c]['in'] for c in self.all_channels]
        self.corner_columns = [x for x in g.frontier(corner_inputs, 'slice', through=lambda i: g.kind(i) != 'interpolated') if g.is_y0_slice(x)]
        self.climate_columns = [x for x in g.frontier(self.climate, 'slice') if g.is_y0_slice(x)]


def ore_vein_rules(rule):
    """The ore_vein rules of a material-rule tree, first occurrence order, equal rules once."""
    out = []

    def walk(r):
        t = r['t']
        if t == 'ref':
            walk(r['rule'])
        elif t == 'sequence':
            for x in r['rules']:
                walk(x)
        elif t == 'condition':
            walk(r['then'])
        elif t == 'ore_vein' and r not in out:
            out.append(r)
    walk(rule)
    return out
