"""IR -> HIP code generator for the TERRAIN step (build time).

Reads the density-function IR (`mccuda-ir/1`) and the terrain model (`mccuda-terrain/1`) and
writes three headers under csrc/gen/:

  model.h       seed-independent constants and tables: the noise layer plan (frequency,
                value factor, smear, octave recipe per layer), MD5 digests of names, the
                flattened climate R-tree, biome and palette data, simplex permutations
  evaluators.h  host+device evaluators, one memoised method per reachable node, for each
                evaluation context of the pipeline
  rules.h       the material-rule tree as inlined functions over a surface context

Evaluation contexts (each has fixed semantics; mixing them changes low bits):

  Columns / ClimateColumns  y = 0 values of the y-independent (sliced) subgraphs, volume path,
                            per distinct 4-block column (corner columns / biome columns)
  Corner                    inputs of the 4x8 interpolation channels at cell corners, volume
                            path, sliced subgraphs read from Columns
  Block / OreBlock          final density and ore density/richness per block; interpolated
                            nodes read the channel values interpolated for the block
  Quart                     the six climate roots at quart positions, volume path
  VolY0                     find_top_surface upper bound, volume path at y = 0
  PointY0 / Point           point path (aquifer queries, ore filler gaps, surface probes);
                            Point delegates y = 0 slices to one shared PointY0 per column

The volume/point split, the compile-time binary-op decisions and the interpolation rules
restate the reference interpreter (cpu/src/volume_eval.cpp:38-674); the context split
follows the GPU pipeline's needs (cuda/codegen/gen_router.py:16-24 names the same contexts).
Branches are lazy: every node is a pure function of position, so skipping a branch a value
does not depend on cannot change it.
"""
import argparse
import json
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import model as M  # noqa: E402


def f32b(bits):
    return struct.unpack('<f', struct.pack('<I', bits & 0xFFFFFFFF))[0]


def f64h(h):
    return struct.unpack('<d', struct.pack('<Q', int(h, 16)))[0]


def fl(v):
    """C++ float literal of an f32 value (exact, hex)."""
    v = M.f32(v)
    if v != v:
        return 'mct::u2f(0x%08Xu)' % struct.unpack('<I', struct.pack('<f', v))[0]
    if v in (float('inf'), float('-inf')):
        return '(%s__builtin_inff())' % ('-' if v < 0 else '')
    return '(%sf)' % v.hex()


def flb(bits):
    return fl(f32b(bits)) if f32b(bits) == f32b(bits) else 'mct::u2f(0x%08Xu)' % bits


def dl(v):
    if v != v:
        return 'mct::u2d(0x%016Xull)' % struct.unpack('<Q', struct.pack('<d', v))[0]
    if v in (float('inf'), float('-inf')):
        return '(%s__builtin_inf())' % ('-' if v < 0 else '')
    return '(%s)' % float(v).hex()


def u64(v):
    return '0x%016Xull' % (v & 0xFFFFFFFFFFFFFFFF)


class IR:
    def __init__(self, doc):
        self.doc = doc
        self.nodes = doc['nodes']
        self.roots = doc['roots']
        for i, n in enumerate(self.nodes):
            if n['id'] != i:
                raise SystemExit('IR node ids must be dense and ordered')

    def children(self, n):
        out = []
        for k in ('in', 'l', 'r', 'alpha', 'first', 'second', 'when_in', 'when_out'):
            if k in n:
                out.append(n[k])
        t = n['t']
        if t == 'noise':
            for a in 'xyz':
                if not n['shift_%s_zero' % a]:
                    out.append(n['shift_' + a])
        elif t == 'interval_select':
            out += n['functions']
        elif t == 'find_top_surface':
            out += [n['density'], n['upper_bound']]
        elif t == 'spline':
            out += spline_coords(n['spline'])
        return list(dict.fromkeys(out))


def spline_coords(s):
    out = []
    if 'coord' in s:
        out.append(s['coord'])
        for v in s['val']:
            out += spline_coords(v)
    return list(dict.fromkeys(out))


class Ctx:
    def __init__(self, name, sem, slices, interp, chans=(), cols=()):
        self.name = name
        self.sem = sem            # 'vol' | 'pt'
        self.slices = slices      # 'col' | 'ident' | 'delegate' | 'subst'
        self.interp = interp      # 'chan' | 'pass' | 'err'
        self.chans = list(chans)
        self.cols = list(cols)


class Gen:
    def __init__(self, ir_doc, tm, ir_hash, tm_hash):
        self.ir = IR(ir_doc)
        self.tm = tm
        self.ir_hash, self.tm_hash = ir_hash, tm_hash
        self.plan = M.LayerPlan(ir_doc)
        self.pre = []             # spline functions and substitution structs, in dependency order
        self.subs = {}
        self.splines = set()
        self.noise_fns = {}
        self.corner_sites = []    # per octave sample site of the corner pass: (layer, xz scale)
        self.corner_site_base = {}
        R = self.ir.roots
        N = self.ir.nodes
        self.final = R['final_density']
        self.ore_rules = []
        self._collect_ores(tm['material_rule'])
        self.ore_roots = []
        for o in self.ore_rules:
            self.ore_roots += [R[o['density']], R[o['richness']]]
        self.climate = [R['climate/' + k] for k in ('temperature', 'vegetation', 'continents', 'erosion', 'depth', 'ridges')]
        fts = [n for n in N if n['t'] == 'find_top_surface']
        if len(fts) != 1:
            raise SystemExit('the IR must have exactly one find_top_surface node')
        self.fts = fts[0]
        a = N[R['aquifer/surface_level']]
        if not (a['t'] == 'slice' and a['axis'] == 'y' and a['coord'] == 0 and a['in'] == self.fts['id']):
            raise SystemExit('aquifer/surface_level must be slice_y0(find_top_surface)')
        c = N[R['surface/chunk_surface_level']]
        i = N[c['in']] if c['t'] == 'slice' else None
        if not (i and i['t'] == 'interpolated' and i['in'] == self.fts['id'] and (i['cell_xz'], i['cell_y']) == (16, 1)):
            raise SystemExit('surface/chunk_surface_level must be slice(interpolated 16x1 (find_top_surface))')
        self.point_roots = [R['aquifer/' + k] for k in ('barrier', 'fluid_level_floodedness', 'fluid_level_spread', 'lava', 'exclusion')]
        self.point_roots += [R[o['filler_gap']] for o in self.ore_rules]
        self.chans = self._leaves([self.final], 'interpolated')
        self.ore_chans = [x for x in self._leaves(self.ore_roots, 'interpolated') if x not in self.chans]
        for x in self.chans + self.ore_chans:
            if (N[x]['cell_xz'], N[x]['cell_y']) != (4, 8):
                raise SystemExit('interpolation channels must be 4x8 cells (node %d)' % x)
        y0 = lambda k: N[k]['axis'] == 'y' and N[k]['coord'] == 0
        self.col_a = [x for x in self._leaves([N[c]['in'] for c in self.chans + self.ore_chans], 'slice', stop=('interpolated',)) if y0(x)]
        self.col_b = [x for x in self._leaves(self.climate, 'slice') if y0(x)]
        self.pt_y0 = [x for x in self._leaves(self.point_roots + [self.fts['density']], 'slice') if y0(x)]

    def _collect_ores(self, r):
        t = r['t']
        if t == 'ref':
            self._collect_ores(r['rule'])
        elif t == 'sequence':
            for x in r['rules']:
                self._collect_ores(x)
        elif t == 'condition':
            self._collect_ores(r['then'])
        elif t == 'ore_vein' and r not in self.ore_rules:
            self.ore_rules.append(r)

    def _leaves(self, roots, kind, stop=()):
        out, seen = [], set()
        stack = list(reversed(roots))
        # DFS in child order; a node of `kind` is a leaf of the walk
        def walk(i):
            if i in seen:
                return
            seen.add(i)
            n = self.ir.nodes[i]
            if n['t'] == kind:
                out.append(i)
                return
            if n['t'] in stop:
                return
            for c in self.ir.children(n):
                walk(c)
        del stack
        for r in roots:
            walk(r)
        return out

    # ---- vanilla compile-time decisions (cpu/src/volume_eval.cpp:38-84) -----------------
    def plan_binary(self, n):
        N = self.ir.nodes
        L, R = N[n['l']], N[n['r']]
        lc, rc = L['t'] == 'constant', R['t'] == 'constant'
        lmin, lmax = f32b(L['range'][0]), f32b(L['range'][1])
        rmin, rmax = f32b(R['range'][0]), f32b(R['range'][1])
        op = n['op']
        if op in ('add', 'sub', 'mul'):
            return op, n['l'], n['r']
        if op == 'div':
            return ('divc' if (not lc and rc) else 'div'), n['l'], n['r']
        if op == 'min':
            if lmax < rmin:
                return 'pass', n['l'], None
            if rmax < lmin:
                return 'pass', n['r'], None
            return ('min', n['r'], n['l']) if lc else ('min', n['l'], n['r'])
        if op == 'max':
            if lmin > rmax:
                return 'pass', n['l'], None
            if rmin > lmax:
                return 'pass', n['r'], None
            return ('max', n['r'], n['l']) if lc else ('max', n['l'], n['r'])
        raise SystemExit('unknown binary op ' + op)

    # ---- noise stack functions ------------------------------------------------------------
    def noise_fn(self, kind, key):
        """kind: 'v' volume (plain octaves), 'g' point get; for blended sets 'sv' / 'sg'.
        key: noise key, or (blended_index, set_name)."""
        name = {'v': 'nv', 'g': 'ng', 'sv': 'bv', 'sg': 'bg'}[kind]
        if isinstance(key, tuple):
            bi, setname = key
            first, count = self.plan.blended[bi][setname]
            name += '_%d_%s' % (bi, setname)
        else:
            k = self.plan.key_index[key]
            first, count = self.plan.noise_range[k]
            name += '_%d' % k
        if name in self.noise_fns:
            return name
        P = self.plan
        L = []
        # Noise stacks are out-of-line calls with the octaves in a loop (not unrolled) over the
        # layer tables: inlined and unrolled, their octave temporaries pile up on the
        # evaluators' memoised values and force register spills and occupancy 1.
        loop = '#pragma clang loop unroll(disable)'
        # On the device the octave table is read through a buffer resource of the wave's first
        # lane's table; lanes of another seed's table take the next pass of the loop (one pass
        # when the wave shares a seed).
        if kind in ('v', 'sv'):
            L.append('MCT_HD MCT_NI float %s(const mct::Octave* O, int32_t bx, int32_t by, int32_t bz, double xz, double ys, float amplitude) {' % name)
            L.append('    float acc = 0.0f;')
            if kind == 'v':
                term = 'acc = acc + (amplitude * kLayerAmp[i]) * mct::perlin_volume(T[i], %s, bx, by, bz, xz * kLayerFreq[i], ys * kLayerFreq[i]);'
            else:
                term = 'acc = acc + (amplitude * kLayerAmp[i]) * mct::smeared_volume(T[i], %s, kLayerSmear[i], bx, by, bz, xz * kLayerFreq[i], ys * kLayerFreq[i]);'
            ret = 'acc'
        else:
            L.append('MCT_HD MCT_NI float %s(const mct::Octave* O, double x, double y, double z) {' % name)
            L.append('    float value = 0.0f;')
            if kind == 'g':
                term = 'value += kLayerAmp[i] * mct::perlin_point(T[i], %s, x * kLayerFreq[i], y * kLayerFreq[i], z * kLayerFreq[i]);'
            else:
                term = 'value += kLayerAmp[i] * mct::smeared_point(T[i], %s, kLayerSmear[i], x * kLayerFreq[i], y * kLayerFreq[i], z * kLayerFreq[i]);'
            ret = 'value'
        L.append('#if defined(__HIP_DEVICE_COMPILE__)')
        L.append('    for (;;) {')
        L.append('        const mct::Octave* T = mct::wave_first(O);')
        L.append('        if (T == O) {')
        L.append('            const __amdgpu_buffer_rsrc_t R = mct::octave_rsrc(T, kNumLayers);')
        L.append(loop)
        L.append('            for (int i = %d; i < %d; ++i) {' % (first, first + count))
        L.append('                ' + term % 'mct::PermBuf(R, i)')
        L.append('            }')
        L.append('            break;')
        L.append('        }')
        L.append('    }')
        L.append('#else')
        L.append('    const mct::Octave* T = O;')
        L.append(loop)
        L.append('    for (int i = %d; i < %d; ++i) {' % (first, first + count))
        L.append('        ' + term % 'mct::PermRef{T[i].p}')
        L.append('    }')
        L.append('#endif')
        L.append('    return %s;' % ret)
        L.append('}')
        self.noise_fns[name] = L
        return name

    def corner_site(self, kind, key, xz):
        """First site of a (noise stack, xz scale) call of the corner pass; its sites are one per
        octave layer, whose x / z records the pass tabulates per block column."""
        if isinstance(key, tuple):
            bi, setname = key
            first, count = self.plan.blended[bi][setname]
        else:
            first, count = self.plan.noise_range[self.plan.key_index[key]]
        ck = (kind, key if not isinstance(key, tuple) else tuple(key), struct.pack('<d', xz))
        if ck not in self.corner_site_base:
            self.corner_site_base[ck] = len(self.corner_sites)
            self.corner_sites += [(first + j, xz) for j in range(count)]
        return self.corner_site_base[ck]

    def noise_fn_tab(self, kind, key):
        """Volume stack of the corner pass: x / z records from the column's LDS table (one per
        layer of the stack, in layer order), y computed here. kind: 'v' plain, 'sv' smeared."""
        name = {'v': 'nvc', 'sv': 'bvc'}[kind]
        if isinstance(key, tuple):
            bi, setname = key
            first, count = self.plan.blended[bi][setname]
            name += '_%d_%s' % (bi, setname)
        else:
            k = self.plan.key_index[key]
            first, count = self.plan.noise_range[k]
            name += '_%d' % k
        if name in self.noise_fns:
            return name
        L = ['MCT_HD MCT_NI float %s(const mct::Octave* O, const MCT_LDS mct::AxisXZ* A, int32_t by, double ys, float amplitude) {' % name,
             '    float acc = 0.0f;']
        if kind == 'v':
            yax = 'const mct::Axis Y = mct::axis_plain(by, ys * kLayerFreq[i], T[i].oy);'
        else:
            yax = 'const mct::Axis Y = mct::axis_smeared(by, ys * kLayerFreq[i], T[i].oy, kLayerSmear[i]);'
        term = 'acc = acc + (amplitude * kLayerAmp[i]) * mct::volume_sample_xz(%s, a, Y);'
        loop = '#pragma clang loop unroll(disable)'
        L += ['#if defined(__HIP_DEVICE_COMPILE__)',
              '    for (;;) {',
              '        const mct::Octave* T = mct::wave_first(O);',
              '        if (T == O) {',
              '            const __amdgpu_buffer_rsrc_t R = mct::octave_rsrc(T, kNumLayers);',
              loop,
              '            for (int i = %d; i < %d; ++i) {' % (first, first + count),
              '                const mct::AxisXZ a = A[i - %d];' % first,
              '                ' + yax,
              '                ' + term % 'mct::PermBuf(R, i)',
              '            }',
              '            break;',
              '        }',
              '    }',
              '#else',
              '    const mct::Octave* T = O;',
              loop,
              '    for (int i = %d; i < %d; ++i) {' % (first, first + count),
              '        const mct::AxisXZ a = A[i - %d];' % first,
              '        ' + yax,
              '        ' + term % 'mct::PermRef{T[i].p}',
              '    }',
              '#endif',
              '    return acc;',
              '}']
        self.noise_fns[name] = L
        return name

    # ---- contexts -------------------------------------------------------------------------
    def _slice_mode(self, n, ctx):
        y0 = n['axis'] == 'y' and n['coord'] == 0
        if y0 and ctx.slices == 'ident':
            return 'ident'
        if y0 and ctx.slices == 'col' and n['id'] in ctx.cols:
            return 'col'
        if y0 and ctx.slices == 'delegate':
            return 'delegate'
        return 'subst'

    def reach(self, ctx, roots):
        out, seen = [], set()

        def walk(i):
            if i in seen:
                return
            seen.add(i)
            n = self.ir.nodes[i]
            t = n['t']
            stop = (t == 'interpolated' and ctx.interp == 'chan') or (t == 'slice' and self._slice_mode(n, ctx) != 'ident') \
                or t == 'find_top_surface'
            if not stop:
                for c in self.ir.children(n):
                    walk(c)
            out.append(i)

        for r in roots:
            walk(r)
        return out

    def sub_struct(self, sem, root):
        key = (sem, root)
        if key in self.subs:
            return self.subs[key]
        name = 'Sub%s%d' % ('V' if sem == 'vol' else 'P', root)
        self.subs[key] = name
        self.pre += self.struct(Ctx(name, sem, 'subst', 'err'), [root])
        return name

    def struct(self, ctx, roots, extra=None):
        reach = self.reach(ctx, roots)
        self.reaches = getattr(self, 'reaches', {})
        self.reaches[ctx.name] = reach
        o = ['struct %s {' % ctx.name, '    const mct::Octave* O;', '    int32_t x, y, z;']
        if ctx.slices == 'col':
            o.append('    const float* col;')
        if ctx.interp == 'chan':
            o.append('    const float* ch;')
        if extra:
            o.append('    ' + extra)
        for i in reach:
            o.append('    float m%d; bool k%d;' % (i, i))
        bodies = []
        for i in reach:
            body = self.node_body(i, ctx)
            bodies.append('    MCT_HD MCT_FI float n%d() {' % i)
            bodies.append('        if (k%d) return m%d;' % (i, i))
            bodies += ['        ' + line for line in body]
            bodies.append('        m%d = v; k%d = true; return v;' % (i, i))
            bodies.append('    }')
        o += bodies
        o.append('};')
        return o

    def node_body(self, i, ctx):
        N = self.ir.nodes
        n = N[i]
        t = n['t']
        vol = ctx.sem == 'vol'
        L = []
        if t == 'interpolated':
            if ctx.interp == 'chan':
                if i not in ctx.chans:
                    raise SystemExit('interpolated node %d is not a channel of %s' % (i, ctx.name))
                L.append('const float v = ch[%d];' % ctx.chans.index(i))
            elif ctx.interp == 'pass':
                if (n['cell_xz'], n['cell_y']) != (4, 8):
                    raise SystemExit('nested interpolation with another cell size (node %d)' % i)
                L.append('const float v = n%d();' % n['in'])
            else:
                raise SystemExit('interpolated node %d cannot be evaluated in %s' % (i, ctx.name))
            return L
        if t == 'slice':
            mode = self._slice_mode(n, ctx)
            if mode == 'ident':
                L.append('const float v = n%d();' % n['in'])
            elif mode == 'col':
                L.append('const float v = col[%d];' % ctx.cols.index(i))
            elif mode == 'delegate':
                L.append('const float v = sub->n%d();' % n['in'])
            else:
                name = self.sub_struct(ctx.sem, n['in'])
                cx = str(n['coord']) if n['axis'] == 'x' else 'x'
                cy = str(n['coord']) if n['axis'] == 'y' else 'y'
                cz = str(n['coord']) if n['axis'] == 'z' else 'z'
                L.append('%s s{O, %s, %s, %s}; const float v = s.n%d();' % (name, cx, cy, cz, n['in']))
            return L
        if t in ('cache', 'blend_density'):
            L.append('const float v = n%d();' % n['in'])
        elif t == 'constant':
            L.append('const float v = %s;' % flb(n['v']))
        elif t == 'simple':
            L.append('const float v = %s;' % ('1.0f' if n['kind'] == 'blend_alpha' else '0.0f'))
        elif t == 'gradient':
            mode = {'clamp_to_edge': 0, 'repeat': 1, 'mirrored_repeat': 2}[n['tiling']]
            frm, to = n['from'], n['to']
            L.append('constexpr float fv = %s, factor = (%s - fv) / static_cast<float>(%d);' % (flb(n['from_value']), flb(n['to_value']), to - frm))
            L.append('const float v = mct::gradient_at(%d, %d, %d, %d, %d, fv, factor, %s);' % (mode, frm, min(frm, to), max(frm, to), to - frm, n['axis']))
        elif t == 'noise':
            xz, ys = dl(f64h(n['xz_scale'])), dl(f64h(n['y_scale']))
            key = n['noise']
            if n['shift_x_zero'] and n['shift_y_zero'] and n['shift_z_zero']:
                if vol and ctx.name == 'Corner':
                    site = self.corner_site('v', key, f64h(n['xz_scale']))
                    L.append('const float v = %s(O, axz + %d, y, %s, 1.0f);' % (self.noise_fn_tab('v', key), site, ys))
                elif vol:
                    L.append('const float v = %s(O, x, y, z, %s, %s, 1.0f);' % (self.noise_fn('v', key), xz, ys))
                else:
                    L.append('const float v = %s(O, static_cast<double>(x) * %s, static_cast<double>(y) * %s, static_cast<double>(z) * %s);'
                             % (self.noise_fn('g', key), xz, ys, xz))
            else:
                L.append('const double nx = static_cast<double>(x) * %s + static_cast<double>(n%d());' % (xz, n['shift_x']))
                if n['shift_y_zero']:
                    L.append('const double ny = static_cast<double>(y) * %s;' % ys)
                else:
                    L.append('const double ny = static_cast<double>(y) * %s + static_cast<double>(n%d());' % (ys, n['shift_y']))
                L.append('const double nz = static_cast<double>(z) * %s + static_cast<double>(n%d());' % (xz, n['shift_z']))
                L.append('const float v = %s(O, nx, ny, nz);' % self.noise_fn('g', key))
        elif t in ('shift_a', 'shift_b', 'shift'):
            key = n['noise']
            if vol:
                if t == 'shift_b':
                    L.append('const float v = %s(O, z, x, 0, 0.25, 0.25, 4.0f);' % self.noise_fn('v', key))
                else:
                    L.append('const float v = %s(O, x, y, z, 0.25, %s, 1.0f) * 4.0f;' % (self.noise_fn('v', key), '0.0' if t == 'shift_a' else '0.25'))
            else:
                g = self.noise_fn('g', key)
                if t == 'shift_b':
                    L.append('const float v = %s(O, static_cast<double>(z) * 0.25, static_cast<double>(x) * 0.25, 0.0) * 4.0f;' % g)
                else:
                    L.append('const float v = %s(O, static_cast<double>(x) * 0.25, static_cast<double>(y) * %s, static_cast<double>(z) * 0.25) * 4.0f;'
                             % (g, '0.0' if t == 'shift_a' else '0.25'))
        elif t == 'blended_noise':
            bi = self.plan.blended_ids.index(i)
            xzm = 684.412 * f64h(n['xz_scale'])
            ym = 684.412 * f64h(n['y_scale'])
            mxz = xzm / f64h(n['xz_factor'])
            my = ym / f64h(n['y_factor'])
            if vol and ctx.name == 'Corner':
                def s(setname, a, b):
                    site = self.corner_site('sv', (bi, setname), a)
                    return '%s(O, axz + %d, y, %s, 1.0f)' % (self.noise_fn_tab('sv', (bi, setname)), site, dl(b))
            elif vol:
                def s(setname, a, b):
                    return '%s(O, x, y, z, %s, %s, 1.0f)' % (self.noise_fn('sv', (bi, setname)), dl(a), dl(b))
            else:
                def s(setname, a, b):
                    return '%s(O, static_cast<double>(x) * %s, static_cast<double>(y) * %s, static_cast<double>(z) * %s)' % (
                        self.noise_fn('sg', (bi, setname)), dl(a), dl(b), dl(a))
            L.append('float a = %s;' % s('main', mxz, my))
            L.append('a = mct::clampf(a + 0.5f, 0.0f, 1.0f);')
            # each limit stack at most once per wave: the lanes that need it call it together
            # (as if / else-if / else, lanes of the three cases would call the stacks up to 4 times)
            L.append('float lo = 0.0f, hi = 0.0f;')
            L.append('if (a != 1.0f) lo = %s;' % s('min_limit', xzm, ym))
            L.append('if (a != 0.0f) hi = %s;' % s('max_limit', xzm, ym))
            L.append('const float v = a == 0.0f ? lo : (a == 1.0f ? hi : mct::lerpf(a, lo, hi));')
        elif t == 'binary':
            kind, a, b = self.plan_binary(n)
            if kind == 'pass':
                L.append('const float v = n%d();' % a)
            elif vol:
                if kind == 'divc':
                    L.append('constexpr float inv = 1.0f / %s;' % flb(N[b]['v']))
                    L.append('const float v = n%d() * inv;' % a)
                else:
                    expr = {'add': 'p + q', 'sub': 'p - q', 'mul': 'p * q', 'div': 'p / q', 'min': 'mct::vmin(p, q)', 'max': 'mct::vmax(p, q)'}[kind]
                    L.append('const float p = n%d(); const float q = n%d(); const float v = %s;' % (a, b, expr))
            else:
                lc = N[n['l']]['t'] == 'constant'
                rc = N[n['r']]['t'] == 'constant'
                if kind in ('add', 'sub'):
                    L.append('const float p = n%d(); const float q = n%d(); const float v = p %s q;' % (n['l'], n['r'], '+' if kind == 'add' else '-'))
                elif kind == 'mul':
                    if lc or rc:
                        L.append('const float p = n%d(); const float q = n%d(); const float v = p * q;' % (n['l'], n['r']))
                    else:
                        L.append('const float p = n%d(); const float v = p == 0.0f ? 0.0f : p * n%d();' % (n['l'], n['r']))
                elif kind == 'divc':
                    L.append('constexpr float inv = 1.0f / %s;' % flb(N[n['r']]['v']))
                    L.append('const float v = n%d() * inv;' % n['l'])
                elif kind == 'div':
                    if lc:
                        L.append('const float v = %s / n%d();' % (flb(N[n['l']]['v']), n['r']))
                    else:
                        L.append('const float p = n%d(); const float v = p == 0.0f ? 0.0f : p / n%d();' % (n['l'], n['r']))
                else:
                    fn = 'mct::jminf' if kind == 'min' else 'mct::jmaxf'
                    if lc or rc:
                        L.append('const float v = %s(n%d(), %s);' % (fn, a, flb(N[b]['v'])))
                    elif kind == 'min':
                        L.append('const float p = n%d(); const float v = p <= %s ? p : mct::jminf(p, n%d());' % (n['l'], flb(N[n['r']]['range'][0]), n['r']))
                    else:
                        L.append('const float p = n%d(); const float v = p >= %s ? p : mct::jmaxf(p, n%d());' % (n['l'], flb(N[n['r']]['range'][1]), n['r']))
        elif t == 'unary':
            f = {'abs': 'mct::jabsf(u)', 'square': 'u * u', 'cube': 'mct::cubef(u)', 'sqrt': 'mct::op_sqrt(u)',
                 'half_negative': 'mct::op_leaky(u, 0.5f)', 'quarter_negative': 'mct::op_leaky(u, 0.25f)',
                 'reciprocal': '1.0f / u', 'negate': '-u', 'squeeze': 'mct::op_squeeze(u)', 'sign': 'mct::op_sign(u)'}.get(n['op'])
            if f is None:
                raise SystemExit('unsupported unary op ' + n['op'])
            L.append('const float u = n%d(); const float v = %s;' % (n['in'], f))
        elif t == 'clamp':
            L.append('const float v = mct::clampf(n%d(), %s, %s);' % (n['in'], flb(n['min']), flb(n['max'])))
        elif t == 'lerp':
            L.append('const float a = n%d();' % n['alpha'])
            L.append('float v;')
            L.append('if (a == 0.0f) v = n%d();' % n['first'])
            L.append('else if (a == 1.0f) v = n%d();' % n['second'])
            L.append('else { const float p = n%d(); const float q = n%d(); v = mct::lerpf(a, p, q); }' % (n['first'], n['second']))
        elif t == 'range_choice':
            L.append('const float c = n%d();' % n['in'])
            L.append('const float v = mct::in_range(c, %s, %s) ? n%d() : n%d();' % (flb(n['min_inclusive']), flb(n['max_exclusive']), n['when_in'], n['when_out']))
        elif t == 'interval_select':
            L.append('const float c = n%d();' % n['in'])
            L.append('float v;')
            fns = n['functions']
            for k, th in enumerate(n['thresholds']):
                L.append('%sif (c < %s) v = n%d();' % ('' if k == 0 else 'else ', flb(th), fns[k]))
            L.append('else v = n%d();' % fns[-1])
        elif t == 'spline':
            coords = spline_coords(n['spline'])
            fname = self.spline_fn(i, coords)
            L.append('const float cs[%d] = {%s};' % (len(coords), ', '.join('n%d()' % c for c in coords)))
            L.append('const float v = %s(cs);' % fname)
        else:
            raise SystemExit('node type %s (node %d) cannot be evaluated in %s' % (t, i, ctx.name))
        return L

    # ---- splines (CubicSpline.Multipoint) ---------------------------------------------------
    def spline_fn(self, node_id, coords):
        name = 'spline%d' % node_id
        if name in self.splines:
            return name
        self.splines.add(name)
        idx = {c: k for k, c in enumerate(coords)}
        L = ['MCT_HD inline float %s(const float* cs) {' % name, '    float v;']
        counter = [0]
        self.emit_spline(self.ir.nodes[node_id]['spline'], 'v', L, idx, counter, '    ')
        L.append('    return v;')
        L.append('}')
        self.pre += L
        return name

    def emit_spline(self, s, target, L, idx, counter, ind):
        if 'k' in s:
            L.append('%s%s = %s;' % (ind, target, flb(s['k'])))
            return
        q = '_%d' % counter[0]
        counter[0] += 1
        loc = [f32b(b) for b in s['loc']]
        der = [f32b(b) for b in s['der']]
        n = len(loc)
        L.append('%s{' % ind)
        i2 = ind + '    '
        L.append('%sconst float in%s = cs[%d];' % (i2, q, idx[s['coord']]))
        # Mth.binarySearch(0, n, i -> in < loc[i]) - 1 == (number of i with !(in < loc[i])) - 1
        L.append('%sconst int st%s = %s - 1;' % (i2, q, ' + '.join('(!(in%s < %s) ? 1 : 0)' % (q, fl(l)) for l in loc)))
        L.append('%sif (st%s < 0) {' % (i2, q))
        L.append('%s    float e%s;' % (i2, q))
        self.emit_spline(s['val'][0], 'e' + q, L, idx, counter, i2 + '    ')
        if der[0] == 0.0:
            L.append('%s    %s = e%s;' % (i2, target, q))
        else:
            L.append('%s    const float dx%s = in%s - %s; %s = e%s + %s * dx%s;' % (i2, q, q, fl(loc[0]), target, q, fl(der[0]), q))
        L.append('%s} else if (st%s >= %d) {' % (i2, q, n - 1))
        L.append('%s    float e%s;' % (i2, q))
        self.emit_spline(s['val'][n - 1], 'e' + q, L, idx, counter, i2 + '    ')
        if der[n - 1] == 0.0:
            L.append('%s    %s = e%s;' % (i2, target, q))
        else:
            L.append('%s    const float dx%s = in%s - %s; %s = e%s + %s * dx%s;' % (i2, q, q, fl(loc[n - 1]), target, q, fl(der[n - 1]), q))
        L.append('%s} else {' % i2)
        L.append('%s    switch (st%s) {' % (i2, q))
        for st in range(n - 1):
            i3 = i2 + '        '
            L.append('%s    case %d: {' % (i2, st))
            L.append('%sconstexpr float x1%s = %s, x2%s = %s, d1%s = %s, d2%s = %s;' % (i3, q, fl(loc[st]), q, fl(loc[st + 1]), q, fl(der[st]), q,
                                                                                 fl(der[st + 1])))
            L.append('%sconst float t%s = (in%s - x1%s) / (x2%s - x1%s);' % (i3, q, q, q, q, q))
            L.append('%sfloat y1%s, y2%s;' % (i3, q, q))
            self.emit_spline(s['val'][st], 'y1' + q, L, idx, counter, i3)
            self.emit_spline(s['val'][st + 1], 'y2' + q, L, idx, counter, i3)
            L.append('%sconst float dx%s = x2%s - x1%s, dy%s = y2%s - y1%s;' % (i3, q, q, q, q, q, q))
            L.append('%sconst float a%s = d1%s * dx%s - dy%s;' % (i3, q, q, q, q))
            L.append('%sconst float b%s = -d2%s * dx%s + dy%s;' % (i3, q, q, q, q))
            L.append('%sconst float l1%s = mct::lerpf(t%s, y1%s, y2%s);' % (i3, q, q, q, q))
            L.append('%sconst float w%s = t%s * (1.0f - t%s);' % (i3, q, q, q))
            L.append('%sconst float l2%s = mct::lerpf(t%s, a%s, b%s);' % (i3, q, q, q, q))
            L.append('%s%s = l1%s + w%s * l2%s;' % (i3, target, q, q, q))
            L.append('%sbreak;' % i3)
            L.append('%s    }' % i2)
        L.append('%s    default: %s = 0.0f; break;' % (i2, target))
        L.append('%s    }' % i2)
        L.append('%s}' % i2)
        L.append('%s}' % ind)

    # ---- material rules -----------------------------------------------------------------
    def emit_rules(self):
        conds, rules = [], []
        self.rand_names = []
        cnt = {'c': 0, 'r': 0}
        key_index = self.plan.key_index

        # 2D noises (at (x, 0, z)) depend on the column only: each gets a slot of the column
        # record the surface kernels compute once per column
        self.col_noise = []

        def col_slot(key):
            g = self.noise_fn('g', key)
            if g not in self.col_noise:
                self.col_noise.append(g)
            return self.col_noise.index(g)

        for k in ('surface', 'surface_secondary', 'clay_bands_offset'):
            col_slot('minecraft:' + k)

        # biomes a column-noise slot can be read under: the union over the rule paths to its
        # conditions of the biome sets those paths require (biome conditions and their
        # negations on the path); slots 0..2 are read outside the rules too
        ALL = (1 << 128) - 1
        slot_biomes = {0: ALL, 1: ALL, 2: ALL}

        def noise_call(key, is3d, bset):
            if not is3d:
                k = col_slot(key)
                slot_biomes[k] = slot_biomes.get(k, 0) | bset
                return 'c.col_noise(%d)' % k
            g = self.noise_fn('g', key)
            return '%s(c.O, static_cast<double>(c.x), static_cast<double>(c.y), static_cast<double>(c.z))' % g

        def biome_set(cn):
            m = 0
            for b in cn['biomes']:
                m |= 1 << b
            return m

        def narrow_biomes(cn, bset):
            while cn['t'] == 'ref':
                cn = cn['cond']
            if cn['t'] == 'biome':
                return bset & biome_set(cn)
            if cn['t'] == 'not':
                inner = cn['cond']
                while inner['t'] == 'ref':
                    inner = inner['cond']
                if inner['t'] == 'biome':
                    return bset & ~biome_set(inner) & ALL
            return bset

        def cond(cn, bset=ALL):
            t = cn['t']
            if t == 'ref':
                return cond(cn['cond'], bset)
            name = 'cond%d' % cnt['c']
            cnt['c'] += 1
            if t == 'biome':
                mask = [0, 0]
                for b in cn['biomes']:
                    if b >= 128:
                        raise SystemExit('biome index >= 128')
                    mask[b >> 6] |= 1 << (b & 63)
                body = 'return c.biome_in(%s, %s);' % (u64(mask[0]), u64(mask[1]))
            elif t == 'noise_threshold':
                if cn['noise'] not in key_index:
                    raise SystemExit('rule noise %s is not in the IR noise list' % cn['noise'])
                body = 'const double v = static_cast<double>(%s); return v >= %s && v <= %s;' % (noise_call(cn['noise'], cn['is_3d'], bset), dl(f64h(cn['min'])),
                                                                                                   dl(f64h(cn['max'])))
            elif t == 'vertical_gradient':
                if cn['random_name'] not in self.rand_names:
                    self.rand_names.append(cn['random_name'])
                body = 'return c.vertical_gradient(%d, %d, %d);' % (self.rand_names.index(cn['random_name']), cn['true_at_and_below'], cn['false_at_and_above'])
            elif t == 'y_above':
                body = 'return c.y + %s >= %d + c.surface_depth * %d;' % ('c.stone_above' if cn['add_stone_depth'] else '0', cn['anchor'],
                                                                           cn['surface_depth_multiplier'])
            elif t == 'water':
                body = 'return c.water_height == INT32_MIN || c.y + %s >= c.water_height + %d + c.surface_depth * %d;' % (
                    'c.stone_above' if cn['add_stone_depth'] else '0', cn['offset'], cn['surface_depth_multiplier'])
            elif t == 'not':
                body = 'return !%s(c);' % cond(cn['cond'], bset)
            elif t == 'stone_depth':
                body = 'return c.stone_depth(%d, %s, %d, %s);' % (cn['offset'], 'true' if cn['add_surface_depth'] else 'false',
                                                                 cn['secondary_depth_range'], 'true' if cn['ceiling'] else 'false')
            elif t == 'temperature':
                body = 'return c.cold_enough_to_snow();'
            elif t == 'steep':
                body = 'return c.grad_x <= -4 || c.grad_z >= 4;'
            elif t == 'hole':
                body = 'return c.surface_depth <= 0;'
            elif t == 'above_preliminary_surface':
                body = 'return c.y >= c.min_surface_level();'
            else:
                raise SystemExit('unsupported condition ' + t)
            conds.append((name, body))
            return name

        # y range of the blocks each ore-vein rule can be reached at: the hull over the rule's
        # paths of the plain y thresholds on them (y_above without stone or surface depth, and
        # its negation); other conditions only narrow it further
        ymin, yend = self.ir.doc['min_y'], self.ir.doc['min_y'] + self.ir.doc['height']
        ore_range = {}

        def deref(cn):
            while cn['t'] == 'ref':
                cn = cn['cond']
            return cn

        def plain_y(cn):
            return cn['t'] == 'y_above' and not cn['add_stone_depth'] and cn['surface_depth_multiplier'] == 0

        def narrow(cn, yr):
            cn = deref(cn)
            if plain_y(cn):
                return (max(yr[0], cn['anchor']), yr[1])
            if cn['t'] == 'not' and plain_y(deref(cn['cond'])):
                return (yr[0], min(yr[1], deref(cn['cond'])['anchor']))
            return yr

        def rule(r, yr=(ymin, yend), bset=ALL):
            t = r['t']
            if t == 'ref':
                return rule(r['rule'], yr, bset)
            name = 'rule%d' % cnt['r']
            cnt['r'] += 1
            if t == 'block':
                body = 'return %d;' % r['state']
            elif t == 'bandlands':
                body = 'return c.band();'
            elif t == 'ore_vein':
                k = self.ore_rules.index(r)
                if yr[0] < yr[1]:
                    lo, hi = ore_range.get(k, yr)
                    ore_range[k] = (min(lo, yr[0]), max(hi, yr[1]))
                body = 'return c.ore_vein(%d);' % k
            elif t == 'condition':
                cn = cond(r['if'], bset)
                th = rule(r['then'], narrow(r['if'], yr), narrow_biomes(r['if'], bset))
                body = 'return %s(c) ? %s(c) : -1;' % (cn, th)
            elif t == 'sequence':
                subs = [rule(x, yr, bset) for x in r['rules']]
                body = 'int s;\n' + '\n'.join('    s = %s(c); if (s >= 0) return s;' % x for x in subs) + '\n    return -1;'
            else:
                raise SystemExit('unsupported rule ' + t)
            rules.append((name, body))
            return name

        root = rule(self.tm['material_rule'])
        o = []
        for name, _ in conds:
            o.append('template <class C> MCT_HD MCT_FI bool %s(C& c);' % name)
        for name, _ in rules:
            o.append('template <class C> MCT_HD MCT_FI int %s(C& c);' % name)
        for name, body in conds:
            o.append('template <class C> MCT_HD MCT_FI bool %s(C& c) { %s }' % (name, body))
        for name, body in rules:
            o.append('template <class C> MCT_HD MCT_FI int %s(C& c) {\n    %s\n}' % (name, body))
        o.append('template <class C> MCT_HD MCT_FI int apply_rules(C& c) { return %s(c); }' % root)
        # per ore-vein rule: the y range its blocks can be in ([lo, hi); empty if unreachable)
        nore = max(len(self.ore_rules), 1)
        rng = [ore_range.get(k, (0, 0)) for k in range(nore)]
        o.append('constexpr int kOreYLo[%d] = {%s};' % (nore, ', '.join(str(a) for a, _ in rng)))
        o.append('constexpr int kOreYHi[%d] = {%s};' % (nore, ', '.join(str(b) for _, b in rng)))
        o.append('constexpr int kColNoises = %d;   // per-column noise slots' % len(self.col_noise))
        o.append('constexpr int kColSurface = 0, kColSecondary = 1, kColClay = 2;')
        # per column-noise slot: the biomes under which the rules can read it (bit b of word
        # b >> 6); a column none of whose candidate biomes is among them never reads the slot
        o.append('constexpr uint64_t kColNoiseBiomes[kColNoises][2] = {%s};' % ', '.join(
            '{%s, %s}' % (u64(slot_biomes.get(k, 0) & ((1 << 64) - 1)), u64(slot_biomes.get(k, 0) >> 64)) for k in range(len(self.col_noise))))
        o.append('MCT_HD MCT_FI float surf_column_noise(const mct::Octave* O, double x, double z, int k) {')
        o.append('    switch (k) {')
        for k, g in enumerate(self.col_noise):
            o.append('        case %d: return %s(O, x, 0.0, z);' % (k, g))
        o.append('    }')
        o.append('    return 0.0f;')
        o.append('}')
        return o

    # ---- model tables --------------------------------------------------------------------
    def model_header(self):
        P, tm, N = self.plan, self.tm, self.ir.nodes
        o = ['// GENERATED by codegen/gen_hip.py from the IR and terrain model. Do not edit.', '#pragma once', '#include <stdint.h>',
             '// Tables the device reads (seed setup) live in constant memory in the device pass.',
             '#if defined(__HIP_DEVICE_COMPILE__)', '#define MCT_TABLE static __constant__ const', '#else', '#define MCT_TABLE static const', '#endif',
             'namespace mctgen {']
        o.append('constexpr uint64_t kIrHash = %s, kTerrainHash = %s;' % (u64(self.ir_hash), u64(self.tm_hash)))
        o.append('constexpr int kMinY = %d, kHeight = %d, kSeaLevel = %d;' % (self.ir.doc['min_y'], self.ir.doc['height'], self.ir.doc['sea_level']))
        if tm['sea_level'] != self.ir.doc['sea_level'] or tm['gen_min_y'] != self.ir.doc['min_y'] or tm['gen_depth'] != self.ir.doc['height']:
            raise SystemExit('IR and terrain model disagree on the world height / sea level')
        o.append('constexpr int kNumLayers = %d, kNumNoises = %d, kNumBlended = %d;' % (P.count, len(P.keys), len(P.blended_ids)))
        o.append('constexpr int kNumChannels = %d, kNumOreChannels = %d, kNumColumns = %d, kNumClimateColumns = %d;' % (
            len(self.chans), len(self.ore_chans), len(self.col_a), len(self.col_b)))
        # first interpolation channel the ore-vein evaluators read (channels below it are
        # not interpolated for ore veins)
        allch = self.chans + self.ore_chans
        used = [allch.index(x) for x in self.reaches.get('OreBlock', []) if x in allch]
        o.append('constexpr int kOreChanFirst = %d;' % (min(used) if used else len(allch)))
        o.append('constexpr int kNumOreRules = %d;' % len(self.ore_rules))
        o.append('constexpr int kFtsLowerBound = %d, kFtsCellHeight = %d;' % (self.fts['lower_bound'], self.fts['cell_height']))
        o.append('static const char* const kNoiseKeys[kNumNoises] = {%s};' % ', '.join('"%s"' % k for k in P.keys))
        o.append('// per layer: kind 0 = NormalNoise (noise key, octave name, fork), 1 = BlendedNoise (node, position in its 40 draws)')
        o.append('struct LayerRecipe { int kind, a, b, c; };')
        o.append('MCT_TABLE LayerRecipe kLayerRecipe[kNumLayers] = {%s};' % ', '.join('{%d, %d, %d, %d}' % r for r in P.recipe))
        o.append('MCT_TABLE double kLayerSmear[kNumLayers] = {%s};' % ', '.join(dl(s) for s in P.smear))
        o.append('MCT_TABLE double kLayerFreq[kNumLayers] = {%s};' % ', '.join(dl(s) for s in P.freq))
        o.append('MCT_TABLE float kLayerAmp[kNumLayers] = {%s};' % ', '.join(fl(s) for s in P.amp))
        o.append('static const int kNoiseFirst[kNumNoises] = {%s};' % ', '.join(str(a) for a, _ in P.noise_range))
        o.append('static const int kNoiseCount[kNumNoises] = {%s};' % ', '.join(str(b) for _, b in P.noise_range))
        md = lambda s: '{%s, %s}' % tuple(u64(v) for v in M.md5_longs(s))
        o.append('struct Md5 { uint64_t lo, hi; };')
        o.append('MCT_TABLE Md5 kNoiseDigest[kNumNoises] = {%s};' % ', '.join(md(k) for k in P.keys))
        o.append('constexpr int kNumOctaveNames = %d;' % len(P.octave_names))
        o.append('MCT_TABLE Md5 kOctaveDigest[kNumOctaveNames] = {%s};' % ', '.join(md(k) for k in P.octave_names))
        o.append('MCT_TABLE Md5 kTerrainDigest = %s, kAquiferDigest = %s, kOreDigest = %s, kClayBandsDigest = %s;' % (
            md('minecraft:terrain'), md('minecraft:aquifer'), md('minecraft:ore'), md('minecraft:clay_bands')))
        o.append('constexpr int kNumRandomNames = %d;' % len(self.rand_names))
        o.append('static const char* const kRandomNames[%d] = {%s};' % (max(1, len(self.rand_names)), ', '.join('"%s"' % k for k in self.rand_names) or '""'))
        o.append('MCT_TABLE Md5 kRandomDigest[%d] = {%s};' % (max(1, len(self.rand_names)), ', '.join(md(k) for k in self.rand_names) or '{0, 0}'))
        # terrain model
        pal = tm['palette']
        if len(pal) > 64:
            raise SystemExit('palette larger than 64 states (u8 output)')
        if len(tm['biome_names']) > 128:
            raise SystemExit('more than 128 biomes')
        flags = []
        for name in pal:
            air = name == 'minecraft:air'
            water = name.startswith('minecraft:water')
            lava = name.startswith('minecraft:lava')
            flags.append((1 if air else 0) | (2 if water else 0) | (4 if lava else 0) | (8 if (water or lava) else 0))
        o.append('constexpr int kPaletteSize = %d;' % len(pal))
        o.append('static const uint8_t kBlockFlags[64] = {%s};' % ', '.join(str(f) for f in flags + [0] * (64 - len(flags))))
        # the same flags as bit masks over palette indices (palette <= 64): tested in registers
        masks = [sum(1 << i for i, f in enumerate(flags) if f & b) for b in (1, 2, 4, 8)]
        o.append('constexpr uint64_t kAirMask = %s, kWaterMask = %s, kLavaMask = %s, kFluidMask = %s;' % tuple(u64(m) for m in masks))

        def blk(name):
            if name not in pal:
                raise SystemExit('palette lacks ' + name)
            return pal.index(name)
        o.append('constexpr int kAir = %d, kWater = %d, kLava = %d, kDefaultBlock = %d, kDefaultFluid = %d, kPackedIce = %d, kSnowBlock = %d;' % (
            blk('minecraft:air'), blk('minecraft:water[level=0]'), blk('minecraft:lava[level=0]'), tm['default_block'], tm['default_fluid'],
            blk('minecraft:packed_ice'), blk('minecraft:snow_block')))
        tc = [blk('minecraft:' + n) for n in ('terracotta', 'white_terracotta', 'orange_terracotta', 'yellow_terracotta', 'brown_terracotta',
                                               'red_terracotta', 'light_gray_terracotta')]
        o.append('static const int kTerracotta[7] = {%s};' % ', '.join(str(x) for x in tc))
        bn = tm['biome_names']
        bmap = {b['name']: b for b in tm['biomes']}
        o.append('constexpr int kNumBiomes = %d;' % len(bn))
        o.append('static const float kBiomeTemp[128] = {%s};' % ', '.join(
            [flb(bmap[n]['temperature']) for n in bn] + ['0.0f'] * (128 - len(bn))))
        o.append('static const uint8_t kBiomeFrozen[128] = {%s};' % ', '.join(
            ['1' if bmap[n]['temperature_modifier'] == 'frozen' else '0' for n in bn] + ['0'] * (128 - len(bn))))
        bix = lambda n: bn.index(n) if n in bn else -1
        o.append('constexpr int kBiomeErodedBadlands = %d, kBiomeFrozenOcean = %d, kBiomeDeepFrozenOcean = %d;' % (
            bix('minecraft:eroded_badlands'), bix('minecraft:frozen_ocean'), bix('minecraft:deep_frozen_ocean')))
        # rule-tree ore veins
        o.append('static const int kOreState[%d][3] = {%s};' % (max(1, len(self.ore_rules)), ', '.join(
            '{%d, %d, %d}' % (r['ore'], r['raw_ore'], r['filler']) for r in self.ore_rules) or '{0, 0, 0}'))
        o.append('static const float kOreRawChance[%d] = {%s};' % (max(1, len(self.ore_rules)), ', '.join(
            flb(r['raw_ore_chance']) for r in self.ore_rules) or '0.0f'))
        # surface noises (MaterialSystem)
        sk = ['surface', 'surface_secondary', 'clay_bands_offset', 'badlands_pillar', 'badlands_pillar_roof', 'badlands_surface',
              'iceberg_pillar', 'iceberg_pillar_roof', 'iceberg_surface']
        for k in sk:
            if 'minecraft:' + k not in P.key_index:
                raise SystemExit('IR lacks the surface noise minecraft:' + k)
        # climate R-tree
        rt = M.RTree(M.load_params(tm))
        mins, maxs, fc, cc, lf = rt.flatten()
        o.append('constexpr int kTreeNodes = %d, kTreeLeaves = %d;' % (len(mins), len(tm['biome_parameters'])))
        o.append('struct TreeNode { int64_t lo[7], hi[7]; int32_t first, count, leaf, pad; };')
        o.append('static const TreeNode kTree[kTreeNodes] = {')
        for k in range(len(mins)):
            o.append('{{%s}, {%s}, %d, %d, %d, 0},' % (','.join('%dll' % v for v in mins[k]), ','.join('%dll' % v for v in maxs[k]), fc[k], cc[k], lf[k]))
        o.append('};')
        o.append('static const uint8_t kLeafBiome[kTreeLeaves] = {%s};' % ','.join(str(p['biome']) for p in tm['biome_parameters']))
        leaf_node = [0] * len(tm['biome_parameters'])
        for k, l in enumerate(lf):
            if l >= 0:
                leaf_node[l] = k
        o.append('static const int32_t kLeafNode[kTreeLeaves] = {%s};' % ','.join(str(v) for v in leaf_node))
        t, fr, info = M.temperature_perms()
        o.append('static const uint8_t kSimplexPerm[5][256] = {%s};' % ', '.join('{%s}' % ','.join(str(v) for v in p) for p in [t] + fr + [info]))
        o.append('}  // namespace mctgen')
        return o

    # ---- entry points -----------------------------------------------------------------------
    def generate(self):
        N = self.ir.nodes
        main = []
        main += self.struct(Ctx('Columns', 'vol', 'ident', 'err'), [N[c]['in'] for c in self.col_a])
        main += self.struct(Ctx('ClimateColumns', 'vol', 'ident', 'err'), [N[c]['in'] for c in self.col_b])
        main += self.struct(Ctx('Corner', 'vol', 'col', 'pass', cols=self.col_a), [N[c]['in'] for c in self.chans + self.ore_chans],
                            extra='const MCT_LDS mct::AxisXZ* axz;')
        main += self.struct(Ctx('Block', 'vol', 'subst', 'chan', chans=self.chans), [self.final])
        main += self.struct(Ctx('OreBlock', 'vol', 'subst', 'chan', chans=self.chans + self.ore_chans), self.ore_roots)
        main += self.struct(Ctx('Quart', 'vol', 'col', 'err', cols=self.col_b), self.climate)
        main += self.struct(Ctx('VolY0', 'vol', 'ident', 'err'), [self.fts['upper_bound']])
        y0_roots = [N[s]['in'] for s in self.pt_y0] + [self.fts['upper_bound']]
        main += self.struct(Ctx('PointY0', 'pt', 'ident', 'err'), y0_roots)
        main += self.struct(Ctx('Point', 'pt', 'delegate', 'err'), self.point_roots + [self.fts['density']], extra='PointY0* sub;')
        rules = self.emit_rules()

        e = []

        def entry(sig, init, body):
            e.append('MCT_HD MCT_FI %s {' % sig)
            e.append('    ' + init)
            e.extend('    ' + b for b in body)
            e.append('}')

        entry('void eval_columns(const mct::Octave* O, int32_t x, int32_t z, float* out)', 'Columns e{O, x, 0, z};',
              ['out[%d] = e.n%d();' % (k, N[c]['in']) for k, c in enumerate(self.col_a)])
        entry('void eval_climate_columns(const mct::Octave* O, int32_t x, int32_t z, float* out)', 'ClimateColumns e{O, x, 0, z};',
              ['out[%d] = e.n%d();' % (k, N[c]['in']) for k, c in enumerate(self.col_b)])
        # x / z records of the corner pass's octave sample sites, per block column
        S = self.corner_sites
        e.append('constexpr int kCornerSites = %d;' % len(S))
        e.append('MCT_TABLE int32_t kCornerSiteLayer[%d] = {%s};' % (max(1, len(S)), ', '.join(str(l) for l, _ in S) or '0'))
        e.append('MCT_TABLE double kCornerSiteScale[%d] = {%s};' % (max(1, len(S)), ', '.join(dl(v) for _, v in S) or '0.0'))
        entry('mct::AxisXZ corner_axes(const mct::Octave* O, int32_t x, int32_t z, int site)',
              'const int i = kCornerSiteLayer[site];', [
                  'const double xs = kCornerSiteScale[site] * kLayerFreq[i];',
                  'mct::Axis X = mct::axis_plain(x, xs, O[i].ox);',
                  'const int xi = X.f & 0xFF;',
                  'X.f = static_cast<int32_t>(O[i].p[xi]) | (static_cast<int32_t>(O[i].p[xi + 1]) << 16);   // perm_x',
                  'return mct::AxisXZ{X, mct::axis_plain(z, xs, O[i].oz)};'])
        entry('void eval_corner(const mct::Octave* O, int32_t x, int32_t y, int32_t z, const float* col, const MCT_LDS mct::AxisXZ* axz, float* out)',
              'Corner e{O, x, y, z, col, axz};',
              ['out[%d] = e.n%d();' % (k, N[c]['in']) for k, c in enumerate(self.chans + self.ore_chans)])
        entry('float eval_block(const mct::Octave* O, int32_t x, int32_t y, int32_t z, const float* ch)', 'Block e{O, x, y, z, ch};',
              ['return e.n%d();' % self.final])
        entry('float eval_ore(const mct::Octave* O, int32_t x, int32_t y, int32_t z, const float* ch, int which)', 'OreBlock e{O, x, y, z, ch};',
              ['switch (which) {'] + ['    case %d: return e.n%d();' % (k, r) for k, r in enumerate(self.ore_roots)] + ['}', 'return 0.0f;'])
        entry('void eval_climate(const mct::Octave* O, int32_t x, int32_t y, int32_t z, const float* col, float* out)', 'Quart e{O, x, y, z, col};',
              ['out[%d] = e.n%d();' % (k, r) for k, r in enumerate(self.climate)])
        # climate values that are a column-record read in the Quart context (a y = 0 slice of the
        # climate columns): the same for every quart of a quart column
        qctx = Ctx('Quart', 'vol', 'col', 'err', cols=self.col_b)
        colconst = 0
        for k, r in enumerate(self.climate):
            n = N[r]
            if n['t'] == 'slice' and self._slice_mode(n, qctx) == 'col':
                colconst |= 1 << k
        e.append('constexpr int kClimateColumnConst = 0x%x;   // bit k: climate value k is the same for all quarts of a quart column' % colconst)
        dens, ub = self.fts['density'], self.fts['upper_bound']
        entry('float fts_probe(const mct::Octave* O, int32_t x, int32_t z, float upper, PointY0& sub)',
              'const float q = upper / static_cast<float>(kFtsCellHeight);', [
                  'const int top = mct::mfloor(static_cast<double>(q)) * kFtsCellHeight;',
                  'if (top <= kFtsLowerBound) return static_cast<float>(kFtsLowerBound);',
                  'for (int probe = top; probe >= kFtsLowerBound; probe -= kFtsCellHeight) {',
                  '    Point e{O, x, probe, z, &sub};',
                  '    if (e.n%d() > 0.0f) return static_cast<float>(probe);' % dens,
                  '}',
                  'return static_cast<float>(kFtsLowerBound);'])
        entry('float fts_volume(const mct::Octave* O, int32_t x, int32_t z)', 'VolY0 u{O, x, 0, z};',
              ['const float ub = u.n%d();' % ub, 'PointY0 sub{O, x, 0, z};', 'return fts_probe(O, x, z, ub, sub);'])
        entry('float fts_point(const mct::Octave* O, int32_t x, int32_t z)', 'PointY0 sub{O, x, 0, z};',
              ['const float ub = sub.n%d();' % ub, 'return fts_probe(O, x, z, ub, sub);'])
        entry('float eval_point(const mct::Octave* O, int32_t x, int32_t y, int32_t z, int which)', 'PointY0 sub{O, x, 0, z};',
              ['Point e{O, x, y, z, &sub};', 'switch (which) {'] + ['    case %d: return e.n%d();' % (k, r) for k, r in enumerate(self.point_roots)] +
              ['}', 'return 0.0f;'])
        # rule noises and surface noise helpers
        sk = {k: self.noise_fn('g', 'minecraft:' + k) for k in ('surface', 'surface_secondary', 'clay_bands_offset', 'badlands_pillar',
                                                                  'badlands_pillar_roof', 'badlands_surface', 'iceberg_pillar',
                                                                  'iceberg_pillar_roof', 'iceberg_surface')}
        for k, fn in sk.items():
            e.append('MCT_HD MCT_FI double surf_%s(const mct::Octave* O, double x, double y, double z) { return static_cast<double>(%s(O, x, y, z)); }' % (k, fn))

        o = ['// GENERATED by codegen/gen_hip.py from the IR and terrain model. Do not edit.', '#pragma once', '#include <stdint.h>',
             '#include "../noise.h"', '#include "model.h"', 'namespace mctgen {']
        for L in self.noise_fns.values():
            o += L
        o += self.pre
        o += main
        o += e
        o.append('}  // namespace mctgen')
        dbg = ['// GENERATED debug dispatchers (bisection against the interpreter); not part of the library.', '#pragma once',
               '#include "evaluators.h"', 'namespace mctgen {']
        dbg.append('inline float debug_corner_node(const mct::Octave* O, int32_t x, int32_t y, int32_t z, const float* col, int id) {')
        dbg.append('    mct::AxisXZ axz[kCornerSites > 0 ? kCornerSites : 1];')
        dbg.append('    for (int k = 0; k < kCornerSites; ++k) axz[k] = corner_axes(O, x, z, k);')
        dbg.append('    Corner e{O, x, y, z, col, axz};')
        dbg.append('    switch (id) {')
        dbg += ['    case %d: return e.n%d();' % (i, i) for i in self.reaches['Corner']]
        dbg.append('    }')
        dbg.append('    return __builtin_nanf("");')
        dbg.append('}')
        dbg.append('inline float debug_point_node(const mct::Octave* O, int32_t x, int32_t y, int32_t z, int id) {')
        dbg.append('    PointY0 sub{O, x, 0, z};')
        dbg.append('    Point e{O, x, y, z, &sub};')
        dbg.append('    switch (id) {')
        dbg += ['    case %d: return e.n%d();' % (i, i) for i in self.reaches['Point']]
        dbg.append('    }')
        dbg.append('    return __builtin_nanf("");')
        dbg.append('}')
        dbg.append('}  // namespace mctgen')
        self.debug_lines = dbg
        r = ['// GENERATED by codegen/gen_hip.py from the terrain model. Do not edit.', '#pragma once', '#include <stdint.h>',
             '#include "evaluators.h"', 'namespace mctgen {']
        r += rules
        r.append('}  // namespace mctgen')
        return o, r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('ir')
    ap.add_argument('terrain')
    ap.add_argument('outdir')
    a = ap.parse_args()
    sys.setrecursionlimit(20000)
    with open(a.ir, encoding='utf-8') as fh:
        ir = json.load(fh)
    with open(a.terrain, encoding='utf-8') as fh:
        tm = json.load(fh)
    if ir.get('format') != 'mccuda-ir/1' or tm.get('format') != 'mccuda-terrain/1':
        raise SystemExit('unexpected input formats')
    g = Gen(ir, tm, M.file_hash(a.ir), M.file_hash(a.terrain))
    ev, rules = g.generate()
    mh = g.model_header()
    os.makedirs(a.outdir, exist_ok=True)
    for name, lines in (('evaluators.h', ev), ('rules.h', rules), ('model.h', mh), ('debug.h', g.debug_lines)):
        path = os.path.join(a.outdir, name)
        text = '\n'.join(lines) + '\n'
        old = open(path).read() if os.path.exists(path) else None
        if old != text:
            with open(path, 'w', encoding='utf-8', newline='\n') as fh:
                fh.write(text)
    print('gen_hip: %d+%d channels, %d+%d column nodes, %d noises, %d layers, %d ore rules, %d point roots -> %s' % (
        len(g.chans), len(g.ore_chans), len(g.col_a), len(g.col_b), len(g.plan.keys), g.plan.count, len(g.ore_rules), len(g.point_roots), a.outdir))


if __name__ == '__main__':
    main()
