"""IR -> HIP compiler of the TERRAIN step (build time).

  python compile_ir.py IR.json TERRAIN.json OUTDIR

Reads the density-function IR (`mccuda-ir/1`) and the terrain model (`mccuda-terrain/1`) and
writes under OUTDIR:

  model.h       seed-independent constants and tables (tables.py)
  evaluators.h  one host + device function per pipeline stage that evaluates the graph
                (placement.py places values into control scopes, lowering.py lowers nodes)
  rules.h       the material-rule tree as one inlined function (rulecomp.py)
  opcounts.h    per-octave operation costs (used by the noise-stack templates' counting
                hooks) and static operation counts per function
  opcounts.json the same static counts for tools

The stage functions and the view each has of the graph:

  eval_columns / eval_climate_columns   y-independent slice inputs at (x, 0, z), volume
                semantics, once per 4-block column (corner columns / biome quart columns)
  eval_corner   the channels' inputs at cell corners; column slices read from the column
                record; noises at the column tabulate their x / z axis records (corner_axes)
  eval_block / eval_ore_*   final density / ore density and richness per block from the
                interpolated channels
  eval_climate  the six climate roots per quart; column slices from the column record
  fts_volume / fts_point    find_top_surface at a column: its upper bound in volume or
                point semantics, then the point-semantics probe loop
  eval_point_*  the point-semantics roots (aquifer noises, ore filler gaps)
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import model as M  # noqa: E402
from densitygraph import DensityGraph, PipelinePlan  # noqa: E402
from lowering import Context, Lowering, NoiseSites, Splines, OCTAVE_OPS, SMEAR_EXTRA, TABLE_RECORD_OPS, dlit  # noqa: E402
from placement import build, render_scope, scope_ops, CLASSES  # noqa: E402
from rulecomp import RuleCompiler  # noqa: E402
from tables import ModelHeader  # noqa: E402

HEADER = '// GENERATED by codegen/compile_ir.py from the IR and terrain model. Do not edit.'
XZ0 = ('x', '0', 'z')


class Compiler:
    def __init__(self, ir, terrain, ir_hash, terrain_hash):
        self.g = DensityGraph(ir)
        self.plan = PipelinePlan(self.g, terrain)
        self.P = M.LayerPlan(ir)
        self.sites = NoiseSites()
        self.splines = Splines()
        self.functions = []          # rendered function text
        self.counts = {}             # function -> {'lower': ops, 'upper': ops}
        self.ir_hash, self.terrain_hash = ir_hash, terrain_hash
        self.ore_channels_read = set()

    def lowering(self, ctx):
        return Lowering(self.g, self.P, ctx, self.sites, self.splines)

    def function(self, name, signature, ctx, fill, ret=None):
        low = self.lowering(ctx)
        b = build(low, fill)
        body = render_scope(b.root, 1)
        self.functions.append('\n'.join(['MCT_HD MCT_FI %s {' % signature] + body + ([] if ret is None else ['    ' + ret]) + ['}']))
        self.counts[name] = {'lower': scope_ops(b.root, False), 'upper': scope_ops(b.root, True)}
        return b

    def outputs(self, ctx, nodes):
        def fill(b, root):
            vals = [b.value((n, ctx.base, ctx.sem), root) for n in nodes]
            for k, v in enumerate(vals):
                root.add('out[%d] = %s;' % (k, v))
        return fill

    def returns(self, ctx, node):
        def fill(b, root):
            root.add('return %s;' % b.value((node, ctx.base, ctx.sem), root))
        return fill

    def dispatcher(self, name, params, args, roots, fn):
        L = ['MCT_HD MCT_FI float %s(%s, int which) {' % (name, params), '    switch (which) {']
        L += ['    case %d: return %s_%d(%s);' % (k, fn, k, args) for k in range(len(roots))]
        L += ['    }', '    return 0.0f;', '}']
        self.functions.append('\n'.join(L))

    def compile(self):
        g, plan = self.g, self.plan
        O = 'const mct::Octave* O'
        cols = Context('Columns', base=XZ0)
        self.function('eval_columns', 'void eval_columns(%s, int32_t x, int32_t z, float* out)' % O, cols,
                      self.outputs(cols, plan.corner_columns))
        ccols = Context('ClimateColumns', base=XZ0)
        self.function('eval_climate_columns', 'void eval_climate_columns(%s, int32_t x, int32_t z, float* out)' % O, ccols,
                      self.outputs(ccols, plan.climate_columns))
        corner = Context('Corner', columns=plan.corner_columns, pass_interpolated=True, tabulated=True)
        self.function('eval_corner', 'void eval_corner(%s, int32_t x, int32_t y, int32_t z, const float* col, const MCT_LDS mct::AxisXZ* axz, const mct::Axis* yt, float* out)' % O,
                      corner, self.outputs(corner, [g.nodes[c]['in'] for c in plan.all_channels]))
        block = Context('Block', channels=plan.channels)
        self.function('eval_block', 'float eval_block(%s, int32_t x, int32_t y, int32_t z, const float* ch)' % O, block,
                      self.returns(block, plan.final))
        ore = Context('OreBlock', channels=plan.all_channels)
        for k, r in enumerate(plan.ore_roots):
            b = self.function('eval_ore_%d' % k, 'float eval_ore_%d(%s, int32_t x, int32_t y, int32_t z, const float* ch)' % (k, O), ore,
                              self.returns(ore, r))
            self.ore_channels_read |= b.lower.channels_read
        self.dispatcher('eval_ore', '%s, int32_t x, int32_t y, int32_t z, const float* ch' % O, 'O, x, y, z, ch', plan.ore_roots, 'eval_ore')
        quart = Context('Quart', columns=plan.climate_columns)
        self.function('eval_climate', 'void eval_climate(%s, int32_t x, int32_t y, int32_t z, const float* col, float* out)' % O, quart,
                      self.outputs(quart, plan.climate))
        low = self.lowering(quart)
        const = 0
        for k, r in enumerate(plan.climate):
            key = (r, quart.base, quart.sem)
            while low.alias(key) is not None:
                key = low.alias(key)
            if low.column_slot(*key) is not None:
                const |= 1 << k
        self.const_lines = ['// bit k: climate value k is a column-record read, the same for every quart of a quart column',
                            'constexpr int kClimateColumnConst = 0x%x;' % const]
        fts = Context('Fts', base=XZ0, sem='point', signs={'pp': '0', 'pn': '(-1)'})
        for name, sem in (('fts_volume', 'volume'), ('fts_point', 'point')):
            self.function(name, 'float %s(%s, int32_t x, int32_t z)' % (name, O), fts, self.probe(sem))
        point = Context('Point', sem='point')
        for k, r in enumerate(plan.point_roots):
            self.function('eval_point_%d' % k, 'float eval_point_%d(%s, int32_t x, int32_t y, int32_t z)' % (k, O), point, self.returns(point, r))
        self.dispatcher('eval_point', '%s, int32_t x, int32_t y, int32_t z' % O, 'O, x, y, z', plan.point_roots, 'eval_point')

    def probe(self, upper_sem):
        """find_top_surface at column (x, z): from the cell at or below the upper bound down to
        the lower bound, the first probe height whose density (point semantics) is positive.
        The probes run as two loops, the heights >= 0 (pp) and then the negative ones (pn): the
        same sequence (the top is a multiple of the cell height, so it passes 0), with the
        sign of the height fixed in each loop, so that noises with a zero y scale are computed
        once per loop instead of once per probe."""
        p = self.plan
        cell, lower = p.fts_cell, p.fts_lower

        def fill(b, root):
            ub = b.value((p.fts_upper, XZ0, upper_sem), root)
            top = b.fresh('top')
            root.add('const int32_t %s = mct::mfloor(static_cast<double>(%s / static_cast<float>(%d))) * %d;' % (top, ub, cell, cell))
            root.add('if (%s <= %d) return static_cast<float>(%d);' % (top, lower, lower))

            def body(var):
                def f(s):
                    d = b.value((p.fts_density, ('x', var, 'z'), 'point'), s)
                    s.add('if (%s > 0.0f) return static_cast<float>(%s);' % (d, var))
                return f
            if lower >= 0:
                b.loop(root, 'for (int32_t pp = %s; pp >= %d; pp -= %d)' % (top, lower, cell), 'pp', body('pp'))
            else:
                b.loop(root, 'for (int32_t pp = %s; pp >= 0; pp -= %d)' % (top, cell), 'pp', body('pp'))
                b.loop(root, 'for (int32_t pn = %s < 0 ? %s : %d; pn >= %d; pn -= %d)' % (top, top, -cell, lower, cell), 'pn', body('pn'))
            root.add('return static_cast<float>(%d);' % lower)
        return fill

    def evaluators(self):
        o = [HEADER, '#pragma once', '#include <stdint.h>', '#include "../stacks.h"', 'namespace mctgen {']
        o += self.splines.lines
        o += self.functions
        o += self.const_lines
        S = self.sites.sites
        o.append('// the corner stage\'s octave sample sites: per site the layer and the xz scale its x / z records use')
        o.append('constexpr int kCornerSites = %d;' % len(S))
        o.append('MCT_TABLE int32_t kCornerSiteLayer[%d] = {%s};' % (max(1, len(S)), ', '.join(str(l) for l, _ in S) or '0'))
        o.append('MCT_TABLE double kCornerSiteScale[%d] = {%s};' % (max(1, len(S)), ', '.join(dlit(v) for _, v in S) or '0.0'))
        o.append('MCT_HD MCT_FI mct::AxisXZ corner_axes(const mct::Octave* O, int32_t x, int32_t z, int site) {')
        o.append('    const int layer = kCornerSiteLayer[site];')
        o.append('    return mct::site_record(O[layer], x, z, kCornerSiteScale[site] * kLayerFreq[layer]);')
        o.append('}')
        Y = self.sites.ysites
        o.append('// the corner stage\'s y sites: per site the layer, the y scale and whether the octave is smeared;')
        o.append('// a seed tabulates each site\'s y axis record at every corner row (kCornerYRows rows, 8 blocks apart)')
        o.append('constexpr int kCornerYSites = %d;' % len(Y))
        o.append('MCT_TABLE int32_t kCornerYLayer[%d] = {%s};' % (max(1, len(Y)), ', '.join(str(l) for l, _, _ in Y) or '0'))
        o.append('MCT_TABLE double kCornerYScale[%d] = {%s};' % (max(1, len(Y)), ', '.join(dlit(v) for _, v, _ in Y) or '0.0'))
        o.append('MCT_TABLE int32_t kCornerYSmeared[%d] = {%s};' % (max(1, len(Y)), ', '.join('1' if sm else '0' for _, _, sm in Y) or '0'))
        o.append('MCT_HD MCT_FI mct::Axis corner_y_axis(const mct::Octave* O, int32_t y, int site) {')
        o.append('    const int layer = kCornerYLayer[site];')
        o.append('    const double s = kCornerYScale[site] * kLayerFreq[layer];')
        o.append('    return kCornerYSmeared[site] ? mct::axis_smeared(y, s, O[layer].oy, kLayerSmear[layer]) : mct::axis_plain(y, s, O[layer].oy);')
        o.append('}')
        for k in ('surface', 'surface_secondary', 'clay_bands_offset', 'badlands_pillar', 'badlands_pillar_roof', 'badlands_surface',
                  'iceberg_pillar', 'iceberg_pillar_roof', 'iceberg_surface'):
            first, count = self.P.noise_range[self.P.key_index['minecraft:' + k]]
            o.append('MCT_HD MCT_FI double surf_%s(const mct::Octave* O, double x, double y, double z) { return static_cast<double>(mct::stack_point<%d, %d>(O, x, y, z)); }'
                     % (k, first, count))
        o.append('}  // namespace mctgen')
        return o

    def opcounts(self):
        o = [HEADER, '#pragma once', 'namespace mctgen {',
             '// operations by class (FP64, FP32, INT32, memory reads), counted at the level of source operations',
             'struct OpCount { int f64, f32, i32, mem; };']
        for name, v in (('kOctaveVolume', OCTAVE_OPS['volume']), ('kOctaveVolumeTab', OCTAVE_OPS['volume_tab']),
                        ('kOctaveVolumeTabY', OCTAVE_OPS['volume_tab_y']), ('kOctavePoint', OCTAVE_OPS['point']),
                        ('kSmearExtra', SMEAR_EXTRA), ('kSiteRecord', TABLE_RECORD_OPS)):
            o.append('constexpr OpCount %s{%d, %d, %d, %d};' % ((name,) + tuple(v)))
        o.append('// static counts per function, without the noise stacks (their octaves count themselves):')
        o.append('// lower = the function\'s own scope and closures, upper = every branch arm once')
        for name, c in self.counts.items():
            o.append('constexpr OpCount kOps_%s_lower{%d, %d, %d, %d}, kOps_%s_upper{%d, %d, %d, %d};' % ((name,) + c['lower'] + (name,) + c['upper']))
        o.append('}  // namespace mctgen')
        js = {'classes': list(CLASSES), 'octave': {'volume': OCTAVE_OPS['volume'], 'volume_tab': OCTAVE_OPS['volume_tab'], 'point': OCTAVE_OPS['point'],
                                                   'smear_extra': SMEAR_EXTRA, 'site_record': TABLE_RECORD_OPS},
              'functions': {k: {'lower': v['lower'], 'upper': v['upper']} for k, v in self.counts.items()}}
        return o, js


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('ir')
    ap.add_argument('terrain')
    ap.add_argument('outdir')
    a = ap.parse_args()
    with open(a.ir, encoding='utf-8') as fh:
        ir = json.load(fh)
    with open(a.terrain, encoding='utf-8') as fh:
        terrain = json.load(fh)
    c = Compiler(ir, terrain, M.file_hash(a.ir), M.file_hash(a.terrain))
    c.compile()
    rules = RuleCompiler(c.plan, c.P)
    rule_lines = rules.compile()
    model = ModelHeader(c.plan, c.P, rules, c.ir_hash, c.terrain_hash).lines(c.ore_channels_read)
    ops_h, ops_js = c.opcounts()
    files = {
        'model.h': [HEADER, '#pragma once', '#include <stdint.h>',
                    '// Tables the device reads while building seed tables live in constant memory in the device pass.',
                    '#if defined(__HIP_DEVICE_COMPILE__)', '#define MCT_TABLE static __constant__ const', '#else', '#define MCT_TABLE static const', '#endif',
                    'namespace mctgen {'] + model + ['}  // namespace mctgen'],
        'evaluators.h': c.evaluators(),
        'rules.h': [HEADER, '#pragma once', '#include <stdint.h>', '#include "evaluators.h"', 'namespace mctgen {'] + rule_lines + ['}  // namespace mctgen'],
        'opcounts.h': ops_h,
    }
    os.makedirs(a.outdir, exist_ok=True)
    for name, lines in files.items():
        write_if_changed(os.path.join(a.outdir, name), '\n'.join(lines) + '\n')
    write_if_changed(os.path.join(a.outdir, 'opcounts.json'), json.dumps(ops_js, indent=1) + '\n')
    stale = os.path.join(a.outdir, 'debug.h')
    if os.path.exists(stale):
        os.remove(stale)
    p = c.plan
    print('compile_ir: %d+%d channels, %d+%d column slots, %d noises, %d layers, %d corner sites, %d ore rules, %d point roots -> %s' % (
        len(p.channels), len(p.ore_channels), len(p.corner_columns), len(p.climate_columns), len(c.P.keys), c.P.count, len(c.sites.sites),
        len(p.ore_rules), len(p.point_roots), a.outdir))


def write_if_changed(path, text):
    old = open(path, encoding='utf-8').read() if os.path.exists(path) else None
    if old != text:
        with open(path, 'w', encoding='utf-8', newline='\n') as fh:
            fh.write(text)


if __name__ == '__main__':
    main()
