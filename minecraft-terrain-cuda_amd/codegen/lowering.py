"""Lowering of density-graph nodes to HIP/C++ expressions inside a placement scope.

A key is (node id, coordinate variant, semantics): the variant is the triple of C++ integer
expressions the node is evaluated at ('x', 'y', 'z' for the block itself; a slice replaces one
of them by its constant; the probe loop of find_top_surface uses its loop variable for y);
the semantics is 'volume' (DensitySampler.sampleVolume: coordinates block * (scale * freq),
keep-first MIN/MAX, no short-circuits) or 'point' (sampleValue: (block * scale) * freq, Java
min/max, zero and range short-circuits). The arithmetic of every node kind restates the
reference interpreter, `cpu/src/volume_eval.cpp:111-328` (volume) and `:357-509` (point),
with the binary decisions of `:38-71`.

A Context says how the stage evaluating a function sees the graph:

  base       the variant of the function's own position
  columns    y = 0 slices read from the stage's column record (`col[k]`) at the base x / z
  channels   interpolated nodes read from the block's interpolated channels (`ch[k]`)
  pass_interpolated   interpolated nodes evaluated in place (the corner stage evaluates the
             channels' inputs at the cell corners)
  tabulated  volume noises at the base x / z read their x / z axis records from the corner
             stage's per-column table (`axz`), computing only y
"""
import struct

from densitygraph import binary_decision, f32_from_bits, f64_from_hex, spline_inputs
from placement import ops, add_ops

# Operation costs of the noise stacks per octave, by class (f64, f32, i32, mem), at the
# level of source operations (one per IEEE operation, conversion, integer operation or table
# read). Per axis of a sample: int -> double, scale multiply, two wrap range compares, offset
# add, floor, double -> int, int -> double, subtract, double -> float (10 FP64) and the
# smoothstep (7 FP32). Lattice hash: 3 masks, 14 permutation reads, 16 index adds. Corners:
# 8 x (3 gradient components: 3 INT; 3 multiplies + 2 adds: 5 FP32); trilinear lerp 7 x 3 FP32;
# the octave's frequency multiplies (FP64) and accumulate (FP32); the octave record reads.
OCTAVE_OPS = {
    # volume path, all three axes computed
    'volume': ops(f64=3 * 10 + 2, f32=3 * 7 + 3 + 8 * 5 + 21 + 3, i32=16 + 24, mem=14 + 5),
    # volume path, x / z from the column table (two records read), y computed
    'volume_tab': ops(f64=10 + 1, f32=7 + 3 + 8 * 5 + 21 + 3, i32=12 + 24, mem=12 + 2 + 4),
    # volume path, x / z from the column table and y from the seed's corner-row table (three
    # records read, no axis computed)
    'volume_tab_y': ops(f32=3 + 8 * 5 + 21 + 3, i32=12 + 24, mem=12 + 2 + 4 + 3),
    # point path: per axis frequency multiply, wrap compares, offset add, floor, two
    # conversions, subtract, to float (9 FP64); fractions minus one, smoothsteps
    'point': ops(f64=3 * 9, f32=3 + 3 * 7 + 8 * 5 + 21 + 2, i32=16 + 24, mem=14 + 5),
}
# smeared octaves add SmearedPerlinNoise.computeFudgeY on y: two compares, divide, add,
# floor, two conversions, multiply, subtract, to float
SMEAR_EXTRA = ops(f64=10)
# one column-table record (corner_axes): two axes of the volume path plus the two first-level
# permutation reads of the x cell
TABLE_RECORD_OPS = ops(f64=2 * 10 + 1, f32=2 * 7, i32=4, mem=2 + 5)


def octave_ops(kind, smeared):
    o = OCTAVE_OPS[kind]
    return add_ops(o, SMEAR_EXTRA) if smeared else o


def flit(v):
    """Exact C++ float literal of an f32 value."""
    v = struct.unpack('<f', struct.pack('<f', v))[0]
    if v != v:
        return 'mct::u2f(0x%08Xu)' % struct.unpack('<I', struct.pack('<f', v))[0]
    if v in (float('inf'), float('-inf')):
        return '(%s__builtin_inff())' % ('-' if v < 0 else '')
    return '(%sf)' % v.hex()


def flit_bits(bits):
    v = f32_from_bits(bits)
    return flit(v) if v == v else 'mct::u2f(0x%08Xu)' % (bits & 0xFFFFFFFF)


def dlit(v):
    if v != v:
        return 'mct::u2d(0x%016Xull)' % struct.unpack('<Q', struct.pack('<d', v))[0]
    if v in (float('inf'), float('-inf')):
        return '(%s__builtin_inf())' % ('-' if v < 0 else '')
    return '(%s)' % float(v).hex()


class Context:
    def __init__(self, name, base=('x', 'y', 'z'), sem='volume', columns=(), channels=(), pass_interpolated=False, tabulated=False,
                 signs=None):
        self.name = name
        # loop variables whose sign is fixed in their loop: name -> an int literal of that sign
        self.signs = signs or {}
        self.base = base
        self.sem = sem
        self.columns = {n: k for k, n in enumerate(columns)}
        self.channels = {n: k for k, n in enumerate(channels)}
        self.pass_interpolated = pass_interpolated
        self.tabulated = tabulated


class NoiseSites:
    """The corner stage's octave sample sites: per (stack, xz scale) one site per layer,
    whose x / z axis records the stage tabulates per block column; and per (stack, y scale,
    smeared) one y site per layer, whose y axis records at the corner rows (the same for every
    column of a seed) the seed setup tabulates."""

    def __init__(self):
        self.sites = []          # (layer, xz scale)
        self.first = {}
        self.ysites = []         # (layer, y scale, smeared)
        self.yfirst = {}

    def site(self, first, count, scale):
        k = (first, count, struct.pack('<d', scale))
        if k not in self.first:
            self.first[k] = len(self.sites)
            self.sites += [(first + j, scale) for j in range(count)]
        return self.first[k]

    def ysite(self, first, count, scale, smeared):
        k = (first, count, struct.pack('<d', scale), bool(smeared))
        if k not in self.yfirst:
            self.yfirst[k] = len(self.ysites)
            self.ysites += [(first + j, scale, bool(smeared)) for j in range(count)]
        return self.yfirst[k]


class Lowering:
    """Emits node values for one context. `plan` is the noise LayerPlan (layer ranges)."""

    def __init__(self, g, layers, ctx, sites, splines):
        self.g, self.P, self.ctx, self.sites, self.splines = g, layers, ctx, sites, splines
        self.channels_read = set()

    # -- keys -------------------------------------------------------------------------------
    def key(self, node, variant=None, sem=None):
        return (node, variant or self.ctx.base, sem or self.ctx.sem)

    def literal(self, key):
        n = self.g.nodes[key[0]]
        if n['t'] == 'constant':
            return flit_bits(n['v'])
        if n['t'] == 'simple':
            return '1.0f' if n['kind'] == 'blend_alpha' else '0.0f'
        return None

    def alias(self, key):
        """The key whose value a pass-through node takes: cache / blend_density markers, slices
        not read from the column record, interpolation evaluated in place, and MIN / MAX nodes
        whose operand ranges decide them at compile time. A point-path noise whose y scale is
        zero reads y only as (double) y * scale, a signed zero fixed by the sign of y: inside a
        loop whose variable has one sign it takes the key at a literal of that sign, which is
        loop-invariant (the same double coordinates, so the same value)."""
        node, var, sem = key
        n = self.g.nodes[node]
        t = n['t']
        if t in ('cache', 'blend_density'):
            return (n['in'], var, sem)
        if sem == 'point' and var[1] in self.ctx.signs and self.zero_y_noise(n):
            return (node, (var[0], self.ctx.signs[var[1]], var[2]), sem)
        if t == 'slice':
            if self.column_slot(node, var, sem) is not None:
                return None
            axis = 'xyz'.index(n['axis'])
            return (n['in'], tuple(str(n['coord']) if a == axis else var[a] for a in range(3)), sem)
        if t == 'interpolated' and not (node in self.ctx.channels and var == self.ctx.base) and self.ctx.pass_interpolated \
                and (n['cell_xz'], n['cell_y']) == (4, 8):
            return (n['in'], var, sem)
        if t == 'binary':
            op, a, _ = binary_decision(self.g, node)
            if op == 'pass':
                return (a, var, sem)
        return None

    @staticmethod
    def zero_y_noise(n):
        """A noise read whose only use of y is (double) y * 0.0 (either zero)."""
        if n['t'] == 'shift_a':
            return True
        return n['t'] == 'noise' and n['shift_x_zero'] and n['shift_y_zero'] and n['shift_z_zero'] and f64_from_hex(n['y_scale']) == 0.0

    def column_slot(self, node, var, sem):
        ctx = self.ctx
        if self.g.is_y0_slice(node) and node in ctx.columns and (var[0], var[2]) == (ctx.base[0], ctx.base[2]) and sem == ctx.sem:
            return ctx.columns[node]
        return None

    def invariant(self, key, var):
        return var not in key[1]

    # -- noise stacks ---------------------------------------------------------------------
    def stack(self, noise_key):
        k = self.P.key_index[noise_key]
        return self.P.noise_range[k]

    def blended_set(self, node, name):
        return self.P.blended[self.P.blended_ids.index(node)][name]

    def stack_call(self, scope, first, count, smeared, sem, coords, xz, ys, amplitude=None):
        """One noise-stack call; coords are C++ int expressions (volume) and the scales f64.
        The stack templates count their own per-octave operations."""
        bx, by, bz = coords
        if sem == 'volume':
            tab = self.ctx.tabulated and (bx, bz) == (self.ctx.base[0], self.ctx.base[2])
            amp = amplitude or '1.0f'
            if tab:
                site = self.sites.site(first, count, xz)
                if by == self.ctx.base[1]:   # at the corner's own y: y records from the seed's table
                    ysite = self.sites.ysite(first, count, ys, smeared)
                    return 'mct::stack_volume_tab_y<%d, %d, kCornerYRows>(O, axz + %d, yt + %d * kCornerYRows, %s)' % (first, count, site, ysite, amp)
                fn = 'mct::stack_volume_tab%s<%d, %d>' % ('_smeared' if smeared else '', first, count)
                return '%s(O, axz + %d, %s, %s, %s)' % (fn, site, by, dlit(ys), amp)
            fn = 'mct::stack_volume%s<%d, %d>' % ('_smeared' if smeared else '', first, count)
            return '%s(O, %s, %s, %s, %s, %s, %s)' % (fn, bx, by, bz, dlit(xz), dlit(ys), amp)
        fn = 'mct::stack_point%s<%d, %d>' % ('_smeared' if smeared else '', first, count)
        return '%s(O, %s, %s, %s)' % (fn, bx, by, bz)

    @staticmethod
    def pscale(c, s):
        """(double) coordinate * scale, a point-path coordinate."""
        return 'static_cast<double>(%s) * %s' % (c, dlit(s))

    # -- emission -----------------------------------------------------------------------------
    def emit(self, b, key, scope):
        node, var, sem = key
        n = self.g.nodes[node]
        t = n['t']
        V = lambda i, v=var, s=sem: b.value((i, v, s), scope)
        vol = sem == 'volume'
        ctx = self.ctx
        if t == 'slice':
            return b.bind(scope, 'col[%d]' % self.column_slot(node, var, sem), ops(mem=1))
        if t == 'interpolated':
            if node in ctx.channels and var == ctx.base:
                self.channels_read.add(ctx.channels[node])
                return b.bind(scope, 'ch[%d]' % ctx.channels[node], ops(mem=1))
            raise SystemExit('interpolated node %d cannot be evaluated in %s at %s' % (node, ctx.name, var))
        if t == 'gradient':
            mode = {'clamp_to_edge': 0, 'repeat': 1, 'mirrored_repeat': 2}[n['tiling']]
            frm, to = n['from'], n['to']
            c = var['xyz'.index(n['axis'])]
            fv, tv = flit_bits(n['from_value']), flit_bits(n['to_value'])
            expr = 'mct::gradient_at(%d, %d, %d, %d, %d, %s, (%s - %s) / static_cast<float>(%d), %s)' % (
                mode, frm, min(frm, to), max(frm, to), to - frm, fv, tv, fv, to - frm, c)
            return b.bind(scope, expr, ops(f32=3, i32=3))
        if t == 'noise':
            return self.emit_noise(b, scope, n, var, sem)
        if t in ('shift_a', 'shift_b', 'shift'):
            first, count = self.stack(n['noise'])
            x, y, z = var
            if vol:
                if t == 'shift_b':
                    return b.bind(scope, self.stack_call(scope, first, count, False, sem, (z, x, '0'), 0.25, 0.25, '4.0f'))
                ys = 0.0 if t == 'shift_a' else 0.25
                return b.bind(scope, self.stack_call(scope, first, count, False, sem, (x, y, z), 0.25, ys) + ' * 4.0f', ops(f32=1))
            if t == 'shift_b':
                c = (self.pscale(z, 0.25), self.pscale(x, 0.25), '0.0')
            else:
                c = (self.pscale(x, 0.25), self.pscale(y, 0.0 if t == 'shift_a' else 0.25), self.pscale(z, 0.25))
            return b.bind(scope, self.stack_call(scope, first, count, False, sem, c, None, None) + ' * 4.0f', ops(f64=6, f32=1))
        if t == 'blended_noise':
            return self.emit_blended(b, scope, node, n, var, sem)
        if t == 'binary':
            return self.emit_binary(b, scope, node, var, sem)
        if t == 'unary':
            u = V(n['in'])
            form, cost = {
                'abs': ('mct::jabsf(%s)', ops(i32=1)), 'square': ('%s * %s', ops(f32=1)), 'cube': ('mct::cubef(%s)', ops(f32=2)),
                'sqrt': ('mct::op_sqrt(%s)', ops(f64=3)), 'half_negative': ('mct::op_leaky(%s, 0.5f)', ops(f32=2)),
                'quarter_negative': ('mct::op_leaky(%s, 0.25f)', ops(f32=2)), 'reciprocal': ('1.0f / %s', ops(f32=1)),
                'negate': ('-%s', ops(f32=1)), 'squeeze': ('mct::op_squeeze(%s)', ops(f32=7)), 'sign': ('mct::op_sign(%s)', ops(f32=3)),
            }.get(n['op'], (None, None))
            if form is None:
                raise SystemExit('unsupported unary op %r (node %d)' % (n['op'], node))
            return b.bind(scope, form.replace('%s', u), cost)
        if t == 'clamp':
            return b.bind(scope, 'mct::clampf(%s, %s, %s)' % (V(n['in']), flit_bits(n['min']), flit_bits(n['max'])), ops(f32=2))
        if t == 'lerp':
            a = V(n['alpha'])
            r = b.fresh()
            scope.add('float %s;' % r)
            b.branch(scope, r, [
                ('if (%s == 0.0f)' % a, lambda s: b.value((n['first'], var, sem), s)),
                ('else if (%s == 1.0f)' % a, lambda s: b.value((n['second'], var, sem), s)),
                ('else', lambda s: self._lerp_arm(b, s, a, n, var, sem)),
            ], ops(f32=2))
            return r
        if t == 'range_choice':
            c = V(n['in'])
            r = b.fresh()
            scope.add('float %s;' % r)
            b.branch(scope, r, [
                ('if (mct::in_range(%s, %s, %s))' % (c, flit_bits(n['min_inclusive']), flit_bits(n['max_exclusive'])),
                 lambda s: b.value((n['when_in'], var, sem), s)),
                ('else', lambda s: b.value((n['when_out'], var, sem), s)),
            ], ops(f32=2))
            return r
        if t == 'interval_select':
            c = V(n['in'])
            r = b.fresh()
            scope.add('float %s;' % r)
            fns, th = n['functions'], n['thresholds']
            arms = [('%sif (%s < %s)' % ('' if k == 0 else 'else ', c, flit_bits(th[k])), lambda s, f=fns[k]: b.value((f, var, sem), s))
                    for k in range(len(th))]
            arms.append(('else', lambda s: b.value((fns[-1], var, sem), s)))
            b.branch(scope, r, arms, ops(f32=len(th)))
            return r
        if t == 'spline':
            coords = spline_inputs(n['spline'])
            vals = [V(c) for c in coords]
            fn = self.splines.function(node, n['spline'], coords)
            arr = b.fresh('cs')
            scope.add('const float %s[%d] = {%s};' % (arr, len(vals), ', '.join(vals)))
            return b.bind(scope, '%s(%s)' % (fn, arr))
        raise SystemExit('node kind %r (node %d) cannot be evaluated in %s' % (t, node, ctx.name))

    def _lerp_arm(self, b, s, a, n, var, sem):
        p = b.value((n['first'], var, sem), s)
        q = b.value((n['second'], var, sem), s)
        return b.bind(s, 'mct::lerpf(%s, %s, %s)' % (a, p, q), ops(f32=3))

    def emit_noise(self, b, scope, n, var, sem):
        first, count = self.stack(n['noise'])
        xz, ys = f64_from_hex(n['xz_scale']), f64_from_hex(n['y_scale'])
        x, y, z = var
        if n['shift_x_zero'] and n['shift_y_zero'] and n['shift_z_zero']:
            if sem == 'volume':
                return b.bind(scope, self.stack_call(scope, first, count, False, sem, var, xz, ys))
            c = (self.pscale(x, xz), self.pscale(y, ys), self.pscale(z, xz))
            return b.bind(scope, self.stack_call(scope, first, count, False, sem, c, None, None), ops(f64=6))
        # shifted noise (always the point formula): (coordinate * scale) + shift, in double
        sh = {a: b.value((n['shift_' + a], var, sem), scope) for a in 'xyz' if not n['shift_%s_zero' % a]}
        nx = '%s + static_cast<double>(%s)' % (self.pscale(x, xz), sh['x'])
        ny = self.pscale(y, ys) if 'y' not in sh else '%s + static_cast<double>(%s)' % (self.pscale(y, ys), sh['y'])
        nz = '%s + static_cast<double>(%s)' % (self.pscale(z, xz), sh['z'])
        cx = b.bind(scope, nx, ops(f64=4), 'const double')
        cy = b.bind(scope, ny, ops(f64=2 if 'y' not in sh else 4), 'const double')
        cz = b.bind(scope, nz, ops(f64=4), 'const double')
        return b.bind(scope, self.stack_call(scope, first, count, False, 'point', (cx, cy, cz), None, None))

    def emit_blended(self, b, scope, node, n, var, sem):
        """BlendedNoise: the main fbm picks the blend of the two 16-octave limit fbms; each
        limit stack runs only when the blend needs it."""
        xzm = 684.412 * f64_from_hex(n['xz_scale'])
        ym = 684.412 * f64_from_hex(n['y_scale'])
        mxz, my = xzm / f64_from_hex(n['xz_factor']), ym / f64_from_hex(n['y_factor'])
        x, y, z = var

        def call(name, sxz, sy):
            first, count = self.blended_set(node, name)
            if sem == 'volume':
                return self.stack_call(scope, first, count, True, sem, var, sxz, sy)
            return self.stack_call(scope, first, count, True, sem, (self.pscale(x, sxz), self.pscale(y, sy), self.pscale(z, sxz)), None, None)
        a = b.fresh('blend')
        lo, hi, v = b.fresh('lo'), b.fresh('hi'), b.fresh()
        pt = ops(f64=6) if sem == 'point' else ops()
        scope.add('float %s = %s;' % (a, call('main', mxz, my)), pt)
        scope.add('%s = mct::clampf(%s + 0.5f, 0.0f, 1.0f);' % (a, a), ops(f32=3))
        # each limit stack at most once per call, and on the device at most once per wave:
        # the lanes needing a stack call it together
        scope.add('float %s = 0.0f, %s = 0.0f;' % (lo, hi))
        scope.add('if (%s != 1.0f) %s = %s;' % (a, lo, call('min_limit', xzm, ym)), pt)
        scope.add('if (%s != 0.0f) %s = %s;' % (a, hi, call('max_limit', xzm, ym)), pt)
        scope.add('const float %s = %s == 0.0f ? %s : (%s == 1.0f ? %s : mct::lerpf(%s, %s, %s));' % (v, a, lo, a, hi, a, lo, hi), ops(f32=5))
        return v

    def emit_binary(self, b, scope, node, var, sem):
        g = self.g
        op, a, c = binary_decision(g, node)
        V = lambda i: b.value((i, var, sem), scope)
        if op == 'divc':
            return b.bind(scope, '%s * (1.0f / %s)' % (V(a), flit_bits(g.nodes[c]['v'])), ops(f32=1))
        if sem == 'volume':
            p, q = V(a), V(c)
            form = {'add': '%s + %s', 'sub': '%s - %s', 'mul': '%s * %s', 'div': '%s / %s', 'min': 'mct::vmin(%s, %s)', 'max': 'mct::vmax(%s, %s)'}[op]
            return b.bind(scope, form % (p, q), ops(f32=1))
        n = g.nodes[node]
        l, r = n['l'], n['r']
        lc, rc = g.is_const(l), g.is_const(r)
        if op in ('add', 'sub'):
            p, q = V(l), V(r)
            return b.bind(scope, '%s %s %s' % (p, '+' if op == 'add' else '-', q), ops(f32=1))
        if op in ('mul', 'div') and not (op == 'mul' and (lc or rc)) and not (op == 'div' and lc):
            # zero short-circuit: the right operand is not evaluated when the left one is 0
            p = V(l)
            res = b.fresh()
            scope.add('float %s;' % res)
            sym = '*' if op == 'mul' else '/'
            b.branch(scope, res, [('if (%s == 0.0f)' % p, lambda s: '0.0f'),
                                  ('else', lambda s: b.bind(s, '%s %s %s' % (p, sym, b.value((r, var, sem), s)), ops(f32=1)))], ops(f32=1))
            return res
        if op == 'mul':
            p, q = V(l), V(r)
            return b.bind(scope, '%s * %s' % (p, q), ops(f32=1))
        if op == 'div':
            return b.bind(scope, '%s / %s' % (flit_bits(g.nodes[l]['v']), V(r)), ops(f32=1))
        fn = 'mct::jminf' if op == 'min' else 'mct::jmaxf'
        if lc or rc:
            return b.bind(scope, '%s(%s, %s)' % (fn, V(a), flit_bits(g.nodes[c]['v'])), ops(f32=1))
        # range short-circuit: the right operand is not evaluated when the left one already
        # decides (left <= min of the right's range for min, >= its max for max)
        p = V(l)
        rb = g.nodes[r]['range']
        test = '%s <= %s' % (p, flit_bits(rb[0])) if op == 'min' else '%s >= %s' % (p, flit_bits(rb[1]))
        res = b.fresh()
        scope.add('float %s;' % res)
        b.branch(scope, res, [('if (%s)' % test, lambda s: p),
                              ('else', lambda s: b.bind(s, '%s(%s, %s)' % (fn, p, b.value((r, var, sem), s)), ops(f32=1)))], ops(f32=1))
        return res


class Splines:
    """CubicSpline.Multipoint functions: one function per (sub)spline, reading its coordinate
    from the evaluator's array of spline inputs. A segment is evaluated by mct::spline_segment
    (the Hermite form vanilla computes), the ends extend linearly by the end derivative."""

    def __init__(self):
        self.lines = []
        self.done = {}
        self.n = 0

    def function(self, node, spline, coords):
        if node in self.done:
            return self.done[node]
        index = {c: k for k, c in enumerate(coords)}
        name = self._emit(node, spline, index)
        self.done[node] = name
        return name

    def _emit(self, node, s, index):
        """Emits the function of spline s (children first); returns its name or a literal."""
        if 'k' in s:
            return flit_bits(s['k'])
        kids = [self._emit(node, v, index) for v in s['val']]
        self.n += 1
        name = 'spline_%d_%d' % (node, self.n)
        loc = [f32_from_bits(x) for x in s['loc']]
        der = [f32_from_bits(x) for x in s['der']]
        m = len(loc)
        val = lambda k: kids[k] if kids[k].startswith('(') or kids[k].startswith('mct::') else '%s(cs)' % kids[k]
        L = ['MCT_HD MCT_FI float %s(const float* cs) {' % name,
             '    const float in = cs[%d];' % index[s['coord']],
             # Mth.binarySearch(0, n, i -> in < loc[i]) - 1: the count of !(in < loc[i]), minus 1
             '    const int at = %s - 1;' % ' + '.join('static_cast<int>(!(in < %s))' % flit(x) for x in loc),
             '    MCT_OPS(0, %d, %d, 0);' % (m, m)]

        def end(k):
            if der[k] == 0.0:
                return val(k)
            return 'mct::spline_extend(in, %s, %s, %s)' % (flit(loc[k]), flit(der[k]), val(k))
        L.append('    if (at < 0) { MCT_OPS(0, %d, 0, 0); return %s; }' % (0 if der[0] == 0.0 else 3, end(0)))
        L.append('    if (at >= %d) { MCT_OPS(0, %d, 0, 0); return %s; }' % (m - 1, 0 if der[m - 1] == 0.0 else 3, end(m - 1)))
        L.append('    MCT_OPS(0, 20, 0, 0);')
        L.append('    switch (at) {')
        for k in range(m - 1):
            L.append('    case %d: return mct::spline_segment(in, %s, %s, %s, %s, %s, %s);' % (
                k, flit(loc[k]), flit(loc[k + 1]), flit(der[k]), flit(der[k + 1]), val(k), val(k + 1)))
        L.append('    default: return 0.0f;')
        L.append('    }')
        L.append('}')
        self.lines += L
        return name
