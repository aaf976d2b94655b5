"""model.h: the seed-independent constants and tables of one IR + terrain model, as C++.

Each entry is (C++ declaration, value) built from the LayerPlan, the R-tree and the terrain
model (model.py computes them). Tables the device reads while building a seed's noise tables
are MCT_TABLE (constant memory in the device pass); the others are plain static arrays the
host copies into device buffers at start-up.
"""
import model as M
from lowering import dlit, flit, flit_bits


def hex64(v):
    return '0x%016Xull' % (v & 0xFFFFFFFFFFFFFFFF)


def md5_init(name):
    lo, hi = M.md5_longs(name)
    return '{%s, %s}' % (hex64(lo), hex64(hi))


def array(values, fmt=str, per_line=16):
    vals = [fmt(v) for v in values]
    if len(vals) <= per_line:
        return '{%s}' % ', '.join(vals)
    rows = [', '.join(vals[i:i + per_line]) for i in range(0, len(vals), per_line)]
    return '{\n    ' + ',\n    '.join(rows) + '}'


class ModelHeader:
    def __init__(self, plan, layers, rules, ir_hash, terrain_hash):
        self.plan, self.P, self.rules = plan, layers, rules
        self.ir_hash, self.terrain_hash = ir_hash, terrain_hash
        self.out = []

    def decl(self, text):
        self.out.append(text)

    def palette_index(self, name):
        pal = self.plan.terrain['palette']
        if name not in pal:
            raise SystemExit('the palette lacks %s' % name)
        return pal.index(name)

    def lines(self, used_channels):
        P, plan, tm, g = self.P, self.plan, self.plan.terrain, self.plan.g
        D = self.decl
        D('constexpr uint64_t kIrHash = %s, kTerrainHash = %s;' % (hex64(self.ir_hash), hex64(self.terrain_hash)))
        D('constexpr int kMinY = %d, kHeight = %d, kSeaLevel = %d;' % (g.min_y, g.height, g.sea_level))
        D('constexpr int kCornerYRows = %d;   // corner rows of the cell grid, 8 blocks apart' % (g.height // 8 + 1))
        # noise layers
        D('constexpr int kNumLayers = %d, kNumNoises = %d, kNumBlended = %d;' % (P.count, len(P.keys), len(P.blended_ids)))
        D('constexpr int kNumChannels = %d, kNumOreChannels = %d, kNumColumns = %d, kNumClimateColumns = %d;' % (
            len(plan.channels), len(plan.ore_channels), len(plan.corner_columns), len(plan.climate_columns)))
        D('// first interpolation channel the ore-vein evaluators read (the lower ones are not interpolated for them)')
        D('constexpr int kOreChanFirst = %d;' % (min(used_channels) if used_channels else len(plan.all_channels)))
        D('constexpr int kNumOreRules = %d;' % len(plan.ore_rules))
        D('constexpr int kFtsLowerBound = %d, kFtsCellHeight = %d;' % (plan.fts_lower, plan.fts_cell))
        D('static const char* const kNoiseKeys[kNumNoises] = %s;' % array(P.keys, lambda k: '"%s"' % k, 4))
        D('// layer recipe: kind 0 = NormalNoise (a noise key, b octave name, c fork), 1 = BlendedNoise (a node, b position in its 40 draws)')
        D('struct LayerRecipe { int kind, a, b, c; };')
        D('MCT_TABLE LayerRecipe kLayerRecipe[kNumLayers] = %s;' % array(P.recipe, lambda r: '{%d, %d, %d, %d}' % r, 8))
        D('MCT_TABLE double kLayerSmear[kNumLayers] = %s;' % array(P.smear, dlit, 6))
        D('MCT_TABLE double kLayerFreq[kNumLayers] = %s;' % array(P.freq, dlit, 6))
        D('MCT_TABLE float kLayerAmp[kNumLayers] = %s;' % array(P.amp, flit, 6))
        D('static const int kNoiseFirst[kNumNoises] = %s;' % array([a for a, _ in P.noise_range]))
        D('static const int kNoiseCount[kNumNoises] = %s;' % array([b for _, b in P.noise_range]))
        # name digests (RandomSupport.seedFromHashOf)
        D('struct Md5 { uint64_t lo, hi; };')
        D('MCT_TABLE Md5 kNoiseDigest[kNumNoises] = %s;' % array(P.keys, md5_init, 2))
        D('constexpr int kNumOctaveNames = %d;' % len(P.octave_names))
        D('MCT_TABLE Md5 kOctaveDigest[kNumOctaveNames] = %s;' % array(P.octave_names, md5_init, 2))
        D('MCT_TABLE Md5 kTerrainDigest = %s, kAquiferDigest = %s, kOreDigest = %s, kClayBandsDigest = %s;' % tuple(
            md5_init('minecraft:' + k) for k in ('terrain', 'aquifer', 'ore', 'clay_bands')))
        names = self.rules.random_names
        D('constexpr int kNumRandomNames = %d;' % len(names))
        D('static const char* const kRandomNames[%d] = %s;' % (max(1, len(names)), array(names or [''], lambda k: '"%s"' % k, 4)))
        D('MCT_TABLE Md5 kRandomDigest[%d] = %s;' % (max(1, len(names)), array(names, md5_init, 2) if names else '{{0, 0}}'))
        self.palette_and_biomes(tm)
        # ore-vein rules of the material rules
        ores = plan.ore_rules
        D('static const int kOreState[%d][3] = %s;' % (max(1, len(ores)), array(ores, lambda r: '{%d, %d, %d}' % (r['ore'], r['raw_ore'], r['filler'])) if ores else '{{0, 0, 0}}'))
        D('static const float kOreRawChance[%d] = %s;' % (max(1, len(ores)), array(ores, lambda r: flit_bits(r['raw_ore_chance'])) if ores else '{0.0f}'))
        for k in ('surface', 'surface_secondary', 'clay_bands_offset', 'badlands_pillar', 'badlands_pillar_roof', 'badlands_surface',
                  'iceberg_pillar', 'iceberg_pillar_roof', 'iceberg_surface'):
            if 'minecraft:' + k not in P.key_index:
                raise SystemExit('the IR lacks the surface noise minecraft:' + k)
        self.tree(tm)
        t, frozen, info = M.temperature_perms()
        D('static const uint8_t kSimplexPerm[5][256] = %s;' % array([t] + frozen + [info], lambda p: array(p, str, 32), 1))
        return self.out

    def palette_and_biomes(self, tm):
        D = self.decl
        pal = tm['palette']
        if len(pal) > 64:
            raise SystemExit('palette larger than 64 states (the output is one byte per block)')
        if len(tm['biome_names']) > 128:
            raise SystemExit('more than 128 biomes')
        flags = []
        for name in pal:
            water, lava = name.startswith('minecraft:water'), name.startswith('minecraft:lava')
            flags.append((1 if name == 'minecraft:air' else 0) | (2 if water else 0) | (4 if lava else 0) | (8 if water or lava else 0))
        D('constexpr int kPaletteSize = %d;' % len(pal))
        D('static const uint8_t kBlockFlags[64] = %s;' % array(flags + [0] * (64 - len(flags))))
        D('// the flags as bit masks over palette indices (air, water, lava, any fluid)')
        D('constexpr uint64_t kAirMask = %s, kWaterMask = %s, kLavaMask = %s, kFluidMask = %s;' % tuple(
            hex64(sum(1 << i for i, f in enumerate(flags) if f & bit)) for bit in (1, 2, 4, 8)))
        ix = self.palette_index
        D('constexpr int kAir = %d, kWater = %d, kLava = %d, kDefaultBlock = %d, kDefaultFluid = %d, kPackedIce = %d, kSnowBlock = %d;' % (
            ix('minecraft:air'), ix('minecraft:water[level=0]'), ix('minecraft:lava[level=0]'), tm['default_block'], tm['default_fluid'],
            ix('minecraft:packed_ice'), ix('minecraft:snow_block')))
        D('static const int kTerracotta[7] = %s;' % array([ix('minecraft:' + n) for n in (
            'terracotta', 'white_terracotta', 'orange_terracotta', 'yellow_terracotta', 'brown_terracotta', 'red_terracotta',
            'light_gray_terracotta')]))
        names = tm['biome_names']
        by_name = {b['name']: b for b in tm['biomes']}
        D('constexpr int kNumBiomes = %d;' % len(names))
        D('static const float kBiomeTemp[128] = %s;' % array([flit_bits(by_name[n]['temperature']) for n in names] + ['0.0f'] * (128 - len(names)),
                                                             str, 8))
        D('static const uint8_t kBiomeFrozen[128] = %s;' % array(
            [1 if by_name[n]['temperature_modifier'] == 'frozen' else 0 for n in names] + [0] * (128 - len(names)), str, 32))
        bix = lambda n: names.index(n) if n in names else -1
        D('constexpr int kBiomeErodedBadlands = %d, kBiomeFrozenOcean = %d, kBiomeDeepFrozenOcean = %d;' % (
            bix('minecraft:eroded_badlands'), bix('minecraft:frozen_ocean'), bix('minecraft:deep_frozen_ocean')))

    def tree(self, tm):
        """The climate R-tree, children of each node contiguous, root first."""
        D = self.decl
        rt = M.RTree(M.load_params(tm))
        mins, maxs, first, count, leaf = rt.flatten()
        nleaves = len(tm['biome_parameters'])
        D('constexpr int kTreeNodes = %d, kTreeLeaves = %d;' % (len(mins), nleaves))
        D('struct TreeNode { int64_t lo[7], hi[7]; int32_t first, count, leaf, pad; };')
        rows = ['{{%s}, {%s}, %d, %d, %d, 0}' % (','.join('%dll' % v for v in mins[k]), ','.join('%dll' % v for v in maxs[k]), first[k], count[k], leaf[k])
                for k in range(len(mins))]
        D('static const TreeNode kTree[kTreeNodes] = {\n' + ',\n'.join(rows) + '};')
        D('static const uint8_t kLeafBiome[kTreeLeaves] = %s;' % array([p['biome'] for p in tm['biome_parameters']], str, 32))
        node_of = [0] * nleaves
        for k, l in enumerate(leaf):
            if l >= 0:
                node_of[l] = k
        D('static const int32_t kLeafNode[kTreeLeaves] = %s;' % array(node_of, str, 16))
