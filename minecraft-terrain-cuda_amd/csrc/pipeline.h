// Per-item bodies of the TERRAIN pipeline, written once for the gfx950 kernels and for the
// host emulation build (the same source compiled by g++), so every stage can be checked on
// a machine without a GPU. A kernel maps its thread index to one call of a body here.
//
// Stages (one batch of chunks, each chunk tagged with a resident seed slot):
//   biomes   climate columns -> per-quart climate target, first minimal R-tree leaf L* and
//            its distance -> tie flags -> last-result chain (per biome chunk)
//   noise    distinct corner columns (dedupe) -> column values -> channel values at the
//            4x8 cell corners
//   aquifer  distinct surface-level cells (dedupe) -> find_top_surface -> per-chunk grid,
//            skip height, aquifer cell locations and fluid status
//   fill     per block: channel interpolation, final density, aquifer substance
//   surface  heights, eroded-badlands / frozen-ocean extensions, material rules
// Reference semantics: cpu/src/terrain.cpp:152-885 (chunk biomes, NoiseBasedAquifer,
// Surface), cpu/src/biome.cpp:201-289 (R-tree search, zoom).
#pragma once

#include <algorithm>
#include <functional>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "jmath.h"
#include "noise.h"
#include "rng.h"
#include "gen/evaluators.h"
#include "gen/rules.h"

namespace mct {

constexpr int kMinY = mctgen::kMinY, kHeight = mctgen::kHeight, kSea = mctgen::kSeaLevel;
constexpr int kQH = kHeight / 4;                   // quart rows
constexpr int kQPer = 16 * kQH;                    // quarts per chunk
constexpr int kPer = 256 * kHeight;                // blocks per chunk
constexpr int kRows = kHeight / 8 + 1;             // corner rows (4x8 cells)
static_assert(kRows == mctgen::kCornerYRows, "the corner-row y tables have one record per corner row");
constexpr int kCh = mctgen::kNumChannels + mctgen::kNumOreChannels;
constexpr int kCols = 25;                          // corner columns per chunk (5 x 5)
constexpr int kGridX = 27, kGridZ = 19, kGridCells = kGridX * kGridZ;   // surface-level quarts per chunk
constexpr int kGridKeys = kGridCells + 4;          // plus the 4 chunk corners of the preliminary surface
constexpr int kWayBelow = -2032 * 16;              // DimensionType.WAY_BELOW_MIN_Y (-2032 << 4)
constexpr int kAqMinGY = -1 + ((kMinY + 1) >= 0 ? (kMinY + 1) / 12 : -((-(kMinY + 1) + 11) / 12));
constexpr int kAqMaxGY = 1 + ((kMinY + kHeight) >= 0 ? (kMinY + kHeight) / 12 : -((-(kMinY + kHeight) + 11) / 12));
constexpr int kAqGY = kAqMaxGY - kAqMinGY + 1;
constexpr int kAqCells = 9 * kAqGY;                // 3 x gridY x 3 per chunk
constexpr int kMaxRandomNames = 8;
constexpr int kMaxOre = 4;
static_assert(mctgen::kNumRandomNames <= kMaxRandomNames, "too many vertical-gradient random names");
static_assert(mctgen::kNumOreRules <= kMaxOre, "too many ore-vein rules");
static_assert(kCh <= 16, "too many interpolation channels");

enum BlockFlag : uint8_t { kFAir = 1, kFWater = 2, kFLava = 4, kFFluid = 8 };

struct ChunkRef {
    int32_t x, z;
    int32_t seed;      // resident seed slot
    int32_t pad;
};

// Per-seed state, built on the device from the seed (seed_layer / seed_blended / seed_misc).
// y axis records of the corner stage's y sites at every corner row: [site][row]
constexpr int kCornerYTab = (mctgen::kCornerYSites > 0 ? mctgen::kCornerYSites : 1) * mctgen::kCornerYRows;

struct SeedDev {
    Octave oct[mctgen::kNumLayers];
    Positional random, aquifer, ore, grad[kMaxRandomNames];
    int64_t zoom_seed;
    int64_t seed;
    uint8_t clay[192];
    Axis ytab[kCornerYTab];
};

struct TreeNodeDev {
    int64_t lo[7], hi[7];
    int32_t first, count, leaf, pad;
};

// The same tree with the bounds as floats (exact while every bound is an integer of magnitude
// at most 2^23), one 64-byte record per node: first child | child count << 24, leaf index.
struct TreeNodeF {
    float lo[7], hi[7];
    int32_t first_count, leaf;
};
constexpr double kTreeExact = 8388608.0;   // 2^23

// Column search groups: for climate coordinate k, the distinct intervals the leaves have on
// it (col_groups[k] of them), or 0 when there are more than kColGroups.
constexpr int kColGroups = 8;

// Seed-independent model tables (uploaded once).
struct ModelDev {
    const TreeNodeDev* tree;
    const TreeNodeF* treef;         // null when some bound is beyond 2^24 (int64 search only)
    const uint8_t* leaf_biome;
    const int32_t* leaf_node;
    const int32_t* leaf_rank;       // position of each leaf in the tree's depth-first order
    int32_t col_groups[7];
    float col_lo[7][kColGroups], col_hi[7][kColGroups];
    uint8_t block_flags[64];
    float biome_temp[128];
    uint8_t biome_frozen[128];
    uint8_t simplex[5][256];       // temperature, frozen x3, biome info
    int32_t terracotta[7];
    int32_t ore_state[kMaxOre][3];
    float ore_raw[kMaxOre];
};

// Levels of internal nodes on the longest root-to-leaf path.
inline int tree_levels(const std::vector<TreeNodeDev>& tree, int n = 0) {
    if (tree.empty() || tree[static_cast<size_t>(n)].leaf >= 0) return 0;
    int deepest = 0;
    for (int k = 0; k < tree[static_cast<size_t>(n)].count; ++k) deepest = std::max(deepest, tree_levels(tree, tree[static_cast<size_t>(n)].first + k));
    return deepest + 1;
}

// Float copy of the tree, or an empty vector when a bound is not exactly representable.
inline std::vector<TreeNodeF> float_tree(const std::vector<TreeNodeDev>& tree) {
    std::vector<TreeNodeF> out(tree.size());
    for (size_t i = 0; i < tree.size(); ++i) {
        for (int k = 0; k < 7; ++k) {
            const int64_t lo = tree[i].lo[k], hi = tree[i].hi[k];
            const int64_t lim = static_cast<int64_t>(kTreeExact);
            if (lo < -lim || lo > lim || hi < -lim || hi > lim || lo > hi) return {};
            out[i].lo[k] = static_cast<float>(lo);
            out[i].hi[k] = static_cast<float>(hi);
        }
        if (tree[i].first >= (1 << 24) || tree[i].count > 127) return {};
        out[i].first_count = tree[i].first | (tree[i].count << 24);
        out[i].leaf = tree[i].leaf;
    }
    return out;
}

// Depth-first rank of every leaf (children in order), by leaf index.
inline std::vector<int32_t> leaf_dfs_rank(const std::vector<TreeNodeDev>& tree, int nleaves) {
    std::vector<int32_t> rank(static_cast<size_t>(nleaves), -1);
    if (tree.empty()) return rank;
    std::vector<std::pair<int, int>> stack{{0, 0}};
    int next = 0;
    if (tree[0].leaf >= 0) { rank[static_cast<size_t>(tree[0].leaf)] = 0; return rank; }
    while (!stack.empty()) {
        auto& top = stack.back();
        const TreeNodeDev& n = tree[static_cast<size_t>(top.first)];
        if (top.second >= n.count) { stack.pop_back(); continue; }
        const int c = n.first + top.second++;
        if (tree[static_cast<size_t>(c)].leaf >= 0) rank[static_cast<size_t>(tree[static_cast<size_t>(c)].leaf)] = next++;
        else stack.push_back({c, 0});
    }
    return rank;
}

// The column search groups of every coordinate (float bounds; only used with the float tree).
inline void column_groups(const std::vector<TreeNodeF>& treef, ModelDev& M) {
    for (int k = 0; k < 7; ++k) {
        std::vector<std::pair<float, float>> g;
        for (const TreeNodeF& n : treef) {
            if (n.leaf < 0) continue;
            const std::pair<float, float> iv{n.lo[k], n.hi[k]};
            if (std::find(g.begin(), g.end(), iv) == g.end()) g.push_back(iv);
            if (g.size() > static_cast<size_t>(kColGroups)) break;
        }
        M.col_groups[k] = (g.empty() || g.size() > static_cast<size_t>(kColGroups)) ? 0 : static_cast<int32_t>(g.size());
        for (int j = 0; j < kColGroups; ++j) {
            M.col_lo[k][j] = j < M.col_groups[k] ? g[static_cast<size_t>(j)].first : 0.0f;
            M.col_hi[k][j] = j < M.col_groups[k] ? g[static_cast<size_t>(j)].second : 0.0f;
        }
    }
}

// ---- keys for the position dedupe -------------------------------------------------------
// A key packs (seed slot, flag, quart x, quart z), each field as narrow as the batch needs, so
// the radix sort runs as few passes as possible:
//   key = (((slot - s0) * 2 + flag) << (bx + bz)) | (ex << bz) | ez.
// ex / ez are the coordinates compressed per axis: the positions a batch keys lie in at most
// kCodecSegs segments along each axis (the union of the chunks' key ranges, the smallest gaps
// merged), and ex is the offset of x in the concatenation of the segments. Batches of a few
// far-apart regions (the border region beside the origin, say) then need the bits of their
// regions' total extent, not of the distance between them.
constexpr int kCodecSegs = 16;
struct AxisCodec {
    int32_t n;                  // segments (1..kCodecSegs)
    int32_t a[kCodecSegs];      // first coordinate of each segment, increasing
    uint32_t o[kCodecSegs];     // compressed offset of each segment, increasing, o[0] = 0
};
struct KeyCodec {
    AxisCodec x, z;
    int32_t s0;
    int32_t bx, bz;
    int32_t bits;     // total key bits
    uint32_t ex, ez;  // exact compressed extents of the two axes
    uint32_t nhi;     // values of the high part, (seed slot - s0) * 2 + flag: 2 x the slot range
};
MCT_HD MCT_FI uint32_t axis_encode(const AxisCodec& c, int32_t q) {
    int j = 0;
    for (int i = 1; i < c.n; ++i) j = q >= c.a[i] ? i : j;
    return c.o[j] + static_cast<uint32_t>(q - c.a[j]);
}
MCT_HD MCT_FI int32_t axis_decode(const AxisCodec& c, uint32_t v) {
    int j = 0;
    for (int i = 1; i < c.n; ++i) j = v >= c.o[i] ? i : j;
    return static_cast<int32_t>(static_cast<int64_t>(c.a[j]) + static_cast<int64_t>(v - c.o[j]));
}
MCT_HD MCT_FI uint64_t make_key(const KeyCodec& k, int32_t slot, int flag, int32_t qx, int32_t qz) {
    const uint64_t hi = static_cast<uint64_t>(slot - k.s0) * 2u + static_cast<uint64_t>(flag);
    return (hi << (k.bx + k.bz)) | (static_cast<uint64_t>(axis_encode(k.x, qx)) << k.bz) | static_cast<uint64_t>(axis_encode(k.z, qz));
}
// A key's index in the dense space [0, nhi * ex * ez) of the codec, the keys' own order
// (high part, then x, then z); for batches whose space is small enough to be marked densely.
MCT_HD MCT_FI uint32_t dense_index(const KeyCodec& k, uint64_t key) {
    const uint32_t ze = static_cast<uint32_t>(key & ((1ull << k.bz) - 1));
    const uint32_t xe = static_cast<uint32_t>((key >> k.bz) & ((1ull << k.bx) - 1));
    const uint32_t hi = static_cast<uint32_t>(key >> (k.bx + k.bz));
    return (hi * k.ex + xe) * k.ez + ze;
}
MCT_HD MCT_FI uint64_t dense_key(const KeyCodec& k, uint32_t d) {
    const uint32_t ze = d % k.ez, r = d / k.ez, xe = r % k.ex, hi = r / k.ex;
    return (static_cast<uint64_t>(hi) << (k.bx + k.bz)) | (static_cast<uint64_t>(xe) << k.bz) | ze;
}
MCT_HD MCT_FI uint64_t dense_size(const KeyCodec& k) { return static_cast<uint64_t>(k.nhi) * k.ex * k.ez; }

MCT_HD MCT_FI void split_key(const KeyCodec& k, uint64_t key, int32_t& slot, int& flag, int32_t& qx, int32_t& qz) {
    qz = axis_decode(k.z, static_cast<uint32_t>(key & ((1ull << k.bz) - 1)));
    qx = axis_decode(k.x, static_cast<uint32_t>((key >> k.bz) & ((1ull << k.bx) - 1)));
    const uint64_t hi = key >> (k.bx + k.bz);
    flag = static_cast<int>(hi & 1u);
    slot = static_cast<int32_t>(hi >> 1) + k.s0;
}

// Codec of a batch (host): per chunk the key range [unit * c + lo, unit * c + hi] along each axis.
// Throws std::range_error when a coordinate leaves int32 or the key 64 bits.
inline int codec_bits(uint64_t range) {
    int b = 0;
    while (b < 64 && (uint64_t{1} << b) < range) ++b;
    return b;
}
inline AxisCodec axis_codec(std::vector<int64_t> c, int unit, int lo, int hi, uint64_t& extent) {
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
    std::vector<std::pair<int64_t, int64_t>> iv;   // disjoint, increasing
    for (int64_t v : c) {
        const int64_t a = unit * v + lo, b = unit * v + hi;
        if (a < INT32_MIN || b > INT32_MAX) throw std::range_error("chunk coordinates outside the range the pipeline's position keys hold");
        if (!iv.empty() && a <= iv.back().second + 1) iv.back().second = std::max(iv.back().second, b);
        else iv.emplace_back(a, b);
    }
    if (static_cast<int>(iv.size()) > kCodecSegs) {   // keep the kCodecSegs - 1 widest gaps
        std::vector<int64_t> gaps;
        for (size_t i = 1; i < iv.size(); ++i) gaps.push_back(iv[i].first - iv[i - 1].second);
        std::vector<int64_t> g = gaps;
        std::nth_element(g.begin(), g.begin() + (kCodecSegs - 2), g.end(), std::greater<int64_t>());
        const int64_t cut = g[kCodecSegs - 2];
        int above = 0;
        for (int64_t x : gaps) above += x > cut ? 1 : 0;
        int ties = kCodecSegs - 1 - above;   // gaps equal to the cut that still split
        std::vector<std::pair<int64_t, int64_t>> m{iv[0]};
        for (size_t i = 1; i < iv.size(); ++i) {
            const int64_t gp = gaps[i - 1];
            const bool split = gp > cut || (gp == cut && ties > 0);
            if (gp == cut && split) --ties;
            if (split) m.push_back(iv[i]);
            else m.back().second = iv[i].second;
        }
        iv.swap(m);
    }
    AxisCodec ac{};
    ac.n = static_cast<int32_t>(iv.size());
    uint64_t off = 0;
    for (int i = 0; i < ac.n; ++i) {
        ac.a[i] = static_cast<int32_t>(iv[static_cast<size_t>(i)].first);
        ac.o[i] = static_cast<uint32_t>(off);
        off += static_cast<uint64_t>(iv[static_cast<size_t>(i)].second - iv[static_cast<size_t>(i)].first + 1);
        if (off > (uint64_t{1} << 32)) throw std::range_error("chunk range too wide for the position keys");
    }
    extent = off;
    return ac;
}
inline KeyCodec make_codec(const ChunkRef* ch, int n, int unit, int xlo, int xhi, int zlo, int zhi) {
    std::vector<int64_t> xs(static_cast<size_t>(n)), zs(static_cast<size_t>(n));
    int64_t mns = INT64_MAX, mxs = INT64_MIN;
    for (int i = 0; i < n; ++i) {
        xs[static_cast<size_t>(i)] = ch[i].x;
        zs[static_cast<size_t>(i)] = ch[i].z;
        mns = std::min<int64_t>(mns, ch[i].seed);
        mxs = std::max<int64_t>(mxs, ch[i].seed);
    }
    if (n == 0) mns = mxs = 0;
    KeyCodec k{};
    uint64_t ex = 1, ez = 1;
    if (n > 0) {
        k.x = axis_codec(xs, unit, xlo, xhi, ex);
        k.z = axis_codec(zs, unit, zlo, zhi, ez);
    } else {
        k.x.n = k.z.n = 1;
    }
    k.s0 = static_cast<int32_t>(mns);
    k.bx = codec_bits(ex);
    k.bz = codec_bits(ez);
    k.ex = static_cast<uint32_t>(std::min<uint64_t>(ex, 0xFFFFFFFFu));   // (a space this wide is never dense)
    k.ez = static_cast<uint32_t>(std::min<uint64_t>(ez, 0xFFFFFFFFu));
    k.nhi = static_cast<uint32_t>(std::min<int64_t>(2 * (mxs - mns + 1), int64_t{1} << 31));
    k.bits = codec_bits(static_cast<uint64_t>(mxs - mns + 1)) + 1 + k.bx + k.bz;
    if (k.bits > 64) throw std::range_error("chunk range too wide for 64-bit position keys");
    return k;
}

// ---- per-seed setup ----------------------------------------------------------------------
// GradientNoise(random, 256): three offsets, then the 256-entry permutation shuffle.
MCT_HD MCT_FI void make_octave(Xoro& r, uint8_t* p, Octave& out) {
    out.ox = next_double(r) * 256.0;
    out.oy = next_double(r) * 256.0;
    out.oz = next_double(r) * 256.0;
    for (int i = 0; i < 256; ++i) p[i] = static_cast<uint8_t>(i);
    for (int i = 0; i < 256; ++i) {
        const int o = next_int(r, 256 - i);
        const uint8_t t = p[i];
        p[i] = p[o + i];
        p[o + i] = t;
    }
    for (int i = 0; i < 512; ++i) out.q[i] = static_cast<uint16_t>(p[i & 255] | p[(i + 1) & 255] << 8);
}

MCT_HD MCT_FI Positional world_random(int64_t seed) {
    Xoro r = xoro_from_seed(seed);
    return fork_positional(r);
}

MCT_HD MCT_FI Digest dg(const mctgen::Md5& m) { return {m.lo, m.hi}; }

// One NormalNoise layer: random.fromHashOf(key) -> first/second positional fork ->
// fromHashOf("octave_N") -> GradientNoise.
MCT_HD inline void seed_layer(SeedDev& S, int64_t seed, int layer, uint8_t* scratch) {
    const mctgen::LayerRecipe& rc = mctgen::kLayerRecipe[layer];
    const Positional P = world_random(seed);
    Xoro r = from_hash(P, dg(mctgen::kNoiseDigest[rc.a]));
    Positional f = fork_positional(r);
    if (rc.c == 1) f = fork_positional(r);
    Xoro o = from_hash(f, dg(mctgen::kOctaveDigest[rc.b]));
    make_octave(o, scratch, S.oct[layer]);
}

// One BlendedNoise node: its min_limit, max_limit and main fbm sets are drawn in sequence
// from random.fromHashOf("minecraft:terrain").
MCT_HD inline void seed_blended(SeedDev& S, int64_t seed, int node, int first_layer, uint8_t* scratch) {
    const Positional P = world_random(seed);
    Xoro r = from_hash(P, dg(mctgen::kTerrainDigest));
    for (int k = 0; k < 40; ++k) make_octave(r, scratch, S.oct[first_layer + k]);
    (void)node;
}

// Positional factories, clay bands (MaterialSystem.generateBands) and the zoom seed.
MCT_HD inline void seed_misc(SeedDev& S, int64_t seed, const ModelDev& M) {
    const Positional P = world_random(seed);
    S.seed = seed;
    S.random = P;
    Xoro a = from_hash(P, dg(mctgen::kAquiferDigest));
    S.aquifer = fork_positional(a);
    Xoro b = from_hash(P, dg(mctgen::kOreDigest));
    S.ore = fork_positional(b);
    for (int k = 0; k < mctgen::kNumRandomNames; ++k) {
        Xoro g = from_hash(P, dg(mctgen::kRandomDigest[k]));
        S.grad[k] = fork_positional(g);
    }
    S.zoom_seed = static_cast<int64_t>(sha256_long_le(seed));
    Xoro r = from_hash(P, dg(mctgen::kClayBandsDigest));
    const int T = M.terracotta[0], WHITE = M.terracotta[1], ORANGE = M.terracotta[2], YELLOW = M.terracotta[3], BROWN = M.terracotta[4],
              RED = M.terracotta[5], LGRAY = M.terracotta[6];
    uint8_t* c = S.clay;
    for (int i = 0; i < 192; ++i) c[i] = static_cast<uint8_t>(T);
    for (int i = 0; i < 192; ++i) {
        i += next_int(r, 5) + 1;
        if (i < 192) c[i] = static_cast<uint8_t>(ORANGE);
    }
    const int widths[3] = {1, 2, 1}, states[3] = {YELLOW, BROWN, RED};
    for (int k = 0; k < 3; ++k) {
        const int count = next_int(r, 15 - 6 + 1) + 6;
        for (int i = 0; i < count; ++i) {
            const int width = widths[k] + next_int(r, 3);
            const int start = next_int(r, 192);
            for (int p = 0; start + p < 192 && p < width; ++p) c[start + p] = static_cast<uint8_t>(states[k]);
        }
    }
    const int whites = next_int(r, 15 - 9 + 1) + 9;
    int i = 0;
    for (int start = 0; i < whites && start < 192; start += next_int(r, 16) + 4) {
        c[start] = static_cast<uint8_t>(WHITE);
        if (start - 1 > 0 && next_bool(r)) c[start - 1] = static_cast<uint8_t>(LGRAY);
        if (start + 1 < 192 && next_bool(r)) c[start + 1] = static_cast<uint8_t>(LGRAY);
        ++i;
    }
}

// Record idx = site * kCornerYRows + row of the seed's corner-row y table (after its octaves).
MCT_HD MCT_FI void seed_yaxis(SeedDev& S, int idx) {
    const int site = idx / mctgen::kCornerYRows, row = idx % mctgen::kCornerYRows;
    S.ytab[idx] = mctgen::corner_y_axis(S.oct, kMinY + row * 8, site);
}

// ---- biomes ------------------------------------------------------------------------------
MCT_HD MCT_FI int64_t quantize(float v) {
    const float q = v * 10000.0f;
    MCT_OPS(0, 4, 0, 0);
    if (q != q) return 0;
    if (q >= 9.2233720368547758e18f) return INT64_MAX;
    if (q <= -9.2233720368547758e18f) return INT64_MIN;
    return static_cast<int64_t>(q);
}

// Squared distance of a target to a parameter box, in Java long arithmetic (wrapping).
MCT_HD MCT_FI int64_t box_distance(const TreeNodeDev& n, const int64_t* t) {
    uint64_t d = 0;
    MCT_OPS(0, 0, 7 * 6, 14);   // per axis: 2 differences, 2 compares, square, sum (64-bit)
    for (int i = 0; i < 7; ++i) {
        const int64_t above = static_cast<int64_t>(static_cast<uint64_t>(t[i]) - static_cast<uint64_t>(n.hi[i]));
        const int64_t below = static_cast<int64_t>(static_cast<uint64_t>(n.lo[i]) - static_cast<uint64_t>(t[i]));
        const int64_t p = above > 0 ? above : (below > 0 ? below : 0);
        d += static_cast<uint64_t>(p) * static_cast<uint64_t>(p);
    }
    return static_cast<int64_t>(d);
}

// First leaf in DFS order at minimal distance (what Climate.RTree.search returns without a
// candidate). Any bound above the minimal distance gives the same leaf, so a neighbouring
// quart's leaf distance + 1 prunes most of the tree. Returns -1 if nothing beats `bound`.
MCT_HD inline int tree_first_min(const TreeNodeDev* T, const int64_t* t, int64_t bound, int64_t& out_d) {
    int stack_node[12], stack_child[12];
    int sp = 0;
    int64_t best = bound;
    int closest = -1;
    if (T[0].leaf >= 0) {
        const int64_t d = box_distance(T[0], t);
        if (best > d) { best = d; closest = 0; }
    } else {
        stack_node[0] = 0;
        stack_child[0] = 0;
        sp = 1;
    }
    while (sp > 0) {
        const TreeNodeDev& n = T[stack_node[sp - 1]];
        const int k = stack_child[sp - 1];
        if (k >= n.count) { --sp; continue; }
        stack_child[sp - 1] = k + 1;
        MCT_OPS(0, 0, 8, 2);
        const int c = n.first + k;
        const int64_t cd = box_distance(T[c], t);
        if (best > cd) {
            if (T[c].leaf >= 0) {
                best = cd;
                closest = c;
            } else if (sp < 12) {
                stack_node[sp] = c;
                stack_child[sp] = 0;
                ++sp;
            }
        }
    }
    out_d = best;
    return closest >= 0 ? T[closest].leaf : -1;
}

// The search over the float tree. With every target coordinate and bound an integer of
// magnitude <= 2^23, each per-axis difference is exact in FP32 (<= 2^24), and as lo <= hi at
// most one of above / below is positive, so max(above, below, 0) is the per-axis distance p;
// p^2 <= 2^48 and a sum of seven < 2^51 are exact in FP64. The distances, their order and the
// leaf found are those of the int64 search above.
MCT_HD MCT_FI double box_distance_f(const TreeNodeF& n, const float* t) {
    double d = 0.0;
    MCT_OPS(7 * 3, 7 * 4, 0, 14);   // per axis: 2 differences and 2 maxima (FP32), widen, square, sum (FP64)
    for (int i = 0; i < 7; ++i) {
        const float p = fmaxf(fmaxf(t[i] - n.hi[i], n.lo[i] - t[i]), 0.0f);
        const double pd = static_cast<double>(p);
        d = d + pd * pd;
    }
    return d;
}

// Distance to the leaf reached by always descending into the nearest child box: an upper
// bound of the minimal leaf distance, to start a search with (+1, so that leaves at exactly the
// minimum are still taken).
MCT_HD inline double tree_greedy_f(const TreeNodeF* T, const TreeNodeF* top, int ntop, const float* t) {
    auto node = [&](int i) -> const TreeNodeF& { return i < ntop ? top[i] : T[i]; };
    int n = 0;
    for (int depth = 0; depth < 16 && node(n).leaf < 0; ++depth) {
        const int fc = node(n).first_count;
        const int first = fc & 0xFFFFFF, count = fc >> 24;
        int pick = first;
        double pd = box_distance_f(node(first), t);
        for (int k = 1; k < count; ++k) {
            const double d = box_distance_f(node(first + k), t);
            if (d < pd) { pd = d; pick = first + k; }
        }
        n = pick;
    }
    return node(n).leaf >= 0 ? box_distance_f(node(n), t) : u2d(0x7FF0000000000000ull);
}

// Nodes below `ntop` are read from `top` (the first nodes of the breadth-first flattening,
// i.e. the upper levels, staged in LDS by the caller), the rest from T.
MCT_HD inline int tree_first_min_f(const TreeNodeF* T, const TreeNodeF* top, int ntop, const float* t, double bound, double& out_d,
                                    int exclude = -1) {
    int stack[12];   // node << 5 | next child (at most 19 children per node)
    int sp = 0;
    double best = bound;
    int closest = -1;
    auto node = [&](int i) -> const TreeNodeF& { return i < ntop ? top[i] : T[i]; };
    if (node(0).leaf >= 0) {
        const double d = box_distance_f(node(0), t);
        if (best > d) { best = d; closest = 0; }
    } else {
        stack[0] = 0;
        sp = 1;
    }
    while (sp > 0) {
        const int e = stack[sp - 1];
        const int fc = node(e >> 5).first_count;
        const int k = e & 31;
        if (k >= (fc >> 24)) { --sp; continue; }
        stack[sp - 1] = e + 1;
        MCT_OPS(1, 0, 6, 1);   // the stack step, the child index, the bound test
        const int c = (fc & 0xFFFFFF) + k;
        const TreeNodeF& cn = node(c);
        const double cd = box_distance_f(cn, t);
        if (best > cd) {
            if (cn.leaf >= 0) {
                if (c != exclude) {
                    best = cd;
                    closest = c;
                }
            } else if (sp < 12) {
                stack[sp++] = c << 5;
            }
        }
    }
    out_d = best;
    return closest >= 0 ? node(closest).leaf : -1;
}

// tree_first_min_f restricted to the leaves whose interval on coordinate k is [glo, ghi]
// (t[k] = glo): subtrees whose interval on k does not cover [glo, ghi] hold none of them and
// are skipped, which leaves the depth-first order of the group's leaves as it is. Returns the
// node index of the first leaf at minimal distance below `bound`, or -1.
MCT_HD inline int tree_first_min_group_f(const TreeNodeF* T, const TreeNodeF* top, int ntop, const float* t, double bound, double& out_d, int k,
                                         float glo, float ghi) {
    int stack[12];
    int sp = 0;
    double best = bound;
    int closest = -1;
    auto node = [&](int i) -> const TreeNodeF& { return i < ntop ? top[i] : T[i]; };
    if (node(0).leaf >= 0) {
        const double d = box_distance_f(node(0), t);
        if (best > d && node(0).lo[k] == glo && node(0).hi[k] == ghi) { best = d; closest = 0; }
    } else {
        stack[0] = 0;
        sp = 1;
    }
    while (sp > 0) {
        const int e = stack[sp - 1];
        const int fc = node(e >> 5).first_count;
        const int j = e & 31;
        if (j >= (fc >> 24)) { --sp; continue; }
        stack[sp - 1] = e + 1;
        MCT_OPS(1, 2, 8, 3);   // the stack step, the child index, the group interval test, the bound test
        const int c = (fc & 0xFFFFFF) + j;
        const TreeNodeF& cn = node(c);
        const bool leaf = cn.leaf >= 0;
        if (leaf ? (cn.lo[k] != glo || cn.hi[k] != ghi) : (cn.lo[k] > glo || cn.hi[k] < ghi)) continue;
        const double cd = box_distance_f(cn, t);
        if (best > cd) {
            if (leaf) {
                best = cd;
                closest = c;
            } else if (sp < 12) {
                stack[sp++] = c << 5;
            }
        }
    }
    out_d = best;
    return closest;
}

// Climate columns of one biome chunk column (y-independent parts at y = 0).
MCT_HD MCT_FI void biome_column(const SeedDev* seeds, const ChunkRef* bch, float* bcol, int idx) {
    const int b = idx >> 4, k = idx & 15;
    const ChunkRef c = bch[b];
    mctgen::eval_climate_columns(seeds[c.seed].oct, c.x * 16 + (k & 3) * 4, c.z * 16 + (k >> 2) * 4, bcol + static_cast<size_t>(idx) * mctgen::kNumClimateColumns);
}

// Quarts per R-tree search thread (a run down one quart column): a launch parameter, 8 unless
// tuned.
constexpr int kBiomeRunDefault = 8;
MCT_HD MCT_FI int biome_runs(int run_len) { return (kQH + run_len - 1) / run_len; }

// Climate target of one quart (idx = biome chunk * kQPer + (qy * 4 + qz) * 4 + qx).
// Climate target coordinates that may vary along a quart column: all but the column-record reads
// the code generator marks and the offset coordinate (always 0).
constexpr int kMayVary = 0x3F & ~mctgen::kClimateColumnConst;
constexpr bool kOneVarying = (kMayVary & (kMayVary - 1)) == 0;
constexpr int kVaryCoord = kMayVary == 0 ? 6 : (kMayVary & 1 ? 0 : kMayVary & 2 ? 1 : kMayVary & 4 ? 2 : kMayVary & 8 ? 3 : kMayVary & 16 ? 4 : 5);

// The climate targets (7 int64 coordinates) of the quarts of a batch's biome chunks, quart
// i = biome chunk * kQPer + (qy * 4 + qz) * 4 + qx. When at most one coordinate can vary along
// a quart column (kOneVarying: the code generator marks the others as column-record reads, the
// same for every quart of the column), a quart column's coordinates are held once (col) and
// each quart's varying one on its own (vary): 8 bytes written per quart instead of 56.
// Otherwise every quart's record is held in full. vary (coordinate kVaryCoord of every quart,
// contiguous) is kept either way: the per-quart final minimum reads that one value.
struct Targets {
    int64_t* full;   // biome chunks x kQPer x 7 (when not kOneVarying)
    int64_t* col;    // biome chunks x 16 x 7: each quart column's coordinates (when kOneVarying)
    int64_t* vary;   // biome chunks x kQPer
    MCT_HD MCT_FI void get(size_t i, int64_t* t) const {
        if constexpr (kOneVarying) {
            const int64_t* c = col + ((i / kQPer) * 16 + (i % kQPer & 15)) * 7;
            for (int k = 0; k < 7; ++k) t[k] = c[k];
            t[kVaryCoord] = vary[i];
        } else {
            for (int k = 0; k < 7; ++k) t[k] = full[i * 7 + k];
        }
    }
    MCT_HD MCT_FI int64_t at(size_t i, int k) const {
        if constexpr (kOneVarying) return k == kVaryCoord ? vary[i] : col[((i / kQPer) * 16 + (i % kQPer & 15)) * 7 + k];
        else return full[i * 7 + k];
    }
};

MCT_HD MCT_FI void biome_target(const SeedDev* seeds, const ChunkRef* bch, const float* bcol, const Targets& T, int idx) {
    const int b = idx / kQPer, q = idx % kQPer;
    const int qx = q & 3, qz = (q >> 2) & 3, qy = q >> 4;
    const ChunkRef c = bch[b];
    const float* cv = bcol + (static_cast<size_t>(b) * 16 + qz * 4 + qx) * mctgen::kNumClimateColumns;
    float clim[6];
    mctgen::eval_climate(seeds[c.seed].oct, c.x * 16 + qx * 4, kMinY + qy * 4, c.z * 16 + qz * 4, cv, clim);
    int64_t t[7];
    for (int k = 0; k < 6; ++k) t[k] = quantize(clim[k]);
    t[6] = 0;
    T.vary[idx] = t[kVaryCoord];
    if constexpr (kOneVarying) {
        if (qy == 0) {
            int64_t* o = T.col + (static_cast<size_t>(b) * 16 + (q & 15)) * 7;
            for (int k = 0; k < 7; ++k) o[k] = t[k];
        }
    } else {
        for (int k = 0; k < 7; ++k) T.full[static_cast<size_t>(idx) * 7 + k] = t[k];
    }
}

// L* and its distance for a run of run_len quarts down one quart column; each search is
// bounded by the previous quart's leaf distance + 1. Targets beyond the float tree's exact
// range take the int64 search here when `wide` is set, else they are marked (L* = -2) for a
// separate pass (biome_search_wide).
MCT_HD inline void biome_search(const ModelDev& M, const TreeNodeF* top, int ntop, const Targets& T, int32_t* lstar, int64_t* mind, int idx,
                                bool wide, int run_len) {
    const int runs = biome_runs(run_len);
    const int b = idx / (16 * runs), rem = idx % (16 * runs);
    const int colk = rem & 15, run = rem >> 4;
    const int qx = colk & 3, qz = colk >> 2;
    const double inf = u2d(0x7FF0000000000000ull);
    int prev = -1;
    const int q1 = (run + 1) * run_len < kQH ? (run + 1) * run_len : kQH;
    for (int qy = run * run_len; qy < q1; ++qy) {
        const size_t i = static_cast<size_t>(b) * kQPer + (qy * 4 + qz) * 4 + qx;
        int64_t t[7];
        T.get(i, t);
        bool small = M.treef != nullptr;
        float td[7];
        for (int k = 0; k < 7; ++k) {
            small = small && t[k] >= -static_cast<int64_t>(kTreeExact) && t[k] <= static_cast<int64_t>(kTreeExact);
            td[k] = static_cast<float>(t[k]);
        }
        if (!small) {
            if (wide) {
                int64_t d = 0;
                lstar[i] = tree_first_min(M.tree, t, INT64_MAX, d);
                mind[i] = d;
            } else {
                lstar[i] = -2;
                mind[i] = -1;
            }
            prev = -1;
            continue;
        }
        double dd = 0.0;
        const double bound = (prev >= 0 ? box_distance_f(M.treef[M.leaf_node[prev]], td) : tree_greedy_f(M.treef, top, ntop, td)) + 1.0;
        int leaf = tree_first_min_f(M.treef, top, ntop, td, bound, dd);
        if (leaf < 0) leaf = tree_first_min_f(M.treef, top, ntop, td, inf, dd);
        lstar[i] = leaf;
        mind[i] = static_cast<int64_t>(dd);
        prev = leaf;
    }
}

// The quarts biome_search marked (targets beyond 2^24): int64 search, unbounded.
MCT_HD inline void biome_search_wide(const ModelDev& M, const Targets& T, int32_t* lstar, int64_t* mind, int idx) {
    if (lstar[idx] != -2) return;
    int64_t d = 0, t[7];
    T.get(static_cast<size_t>(idx), t);
    lstar[idx] = tree_first_min(M.tree, t, INT64_MAX, d);
    mind[idx] = d;
}

// L* and its distance for all quarts of one quart column (idx = biome chunk * 16 + qz * 4 + qx).
// Along a column usually one climate coordinate varies (the depth) and the leaves take a few
// distinct intervals on it (groups). A leaf's distance is then A (over the other coordinates,
// fixed for the column) plus the squared distance of the column's value v on that coordinate
// to the leaf's interval, the same for its whole group. One search per group restricted to
// the group (at v = the group's lower bound, where that term is 0) gives the group's minimal
// A and its first leaf in depth-first order; a quart's L* is the leaf of the group minimising
// A_g + d_g(v)^2, ties to the leaf first in depth-first order, which is the leaf and distance
// the unrestricted search gives. Columns that vary in more than one coordinate, or in one with
// too many groups, or whose targets leave the float tree's exact range, take the per-quart
// search.
MCT_HD inline void biome_search_column(const ModelDev& M, const TreeNodeF* top, int ntop, const Targets& T, int32_t* lstar, int64_t* mind, int idx) {
    const int b = idx >> 4, qx = idx & 3, qz = (idx >> 2) & 3;
    const size_t col = static_cast<size_t>(b) * kQPer + qz * 4 + qx;   // + qy * 16: the column's quarts
    int64_t t0[7];
    T.get(col, t0);
    auto in_range = [](int64_t v) { return v >= -static_cast<int64_t>(kTreeExact) && v <= static_cast<int64_t>(kTreeExact); };
    // The code generator marks the climate values that are column-record reads (the same for
    // every quart of the column); the target's offset coordinate is always 0. When at most one
    // coordinate is left, the column varies in that one alone and needs no scan of its targets;
    // a quart whose varying value leaves the float tree's exact range is marked for the int64
    // search (L* = -2) as in the per-quart search.
    int vary = 0, k = -1;
    bool small = M.treef != nullptr, scanned = false;
    if constexpr ((kMayVary & (kMayVary - 1)) == 0) {
        for (int j = 0; j < 7; ++j)
            if (!((kMayVary >> j) & 1)) small = small && in_range(t0[j]);
        vary = kMayVary;
    } else {
        scanned = true;
        for (int qy = 0; qy < kQH; ++qy) {
            int64_t t[7];
            T.get(col + static_cast<size_t>(qy) * 16, t);
            for (int j = 0; j < 7; ++j) {
                if (t[j] != t0[j]) vary |= 1 << j;
                small = small && in_range(t[j]);
            }
        }
    }
    if (small) {
        if (vary == 0) {
            for (int j = 6; j >= 0; --j)
                if (M.col_groups[j] > 0) k = j;
        } else if ((vary & (vary - 1)) == 0) {
            const int j = __builtin_ctz(static_cast<unsigned>(vary));
            if (M.col_groups[j] > 0) k = j;
        }
    }
    if (k < 0) {
        biome_search(M, top, ntop, T, lstar, mind, idx, false, kQH);
        return;
    }
    auto node = [&](int i) -> const TreeNodeF& { return i < ntop ? top[i] : M.treef[i]; };
    const double inf = u2d(0x7FF0000000000000ull);
    float tf[7];
    for (int j = 0; j < 7; ++j) tf[j] = static_cast<float>(t0[j]);
    const int G = M.col_groups[k];
    double A[kColGroups];
    int leaf[kColGroups], rank[kColGroups];
    float lo[kColGroups], hi[kColGroups];
#pragma unroll
    for (int g = 0; g < kColGroups; ++g) {
        leaf[g] = -1;
        rank[g] = INT32_MAX;
        A[g] = inf;
        lo[g] = M.col_lo[k][g];
        hi[g] = M.col_hi[k][g];
        if (g < G) {
            float tg[7];
            for (int j = 0; j < 7; ++j) tg[j] = j == k ? lo[g] : tf[j];
            double d = 0.0;
            int n = tree_first_min_group_f(M.treef, top, ntop, tg, tree_greedy_f(M.treef, top, ntop, tg) + 1.0, d, k, lo[g], hi[g]);
            if (n < 0) n = tree_first_min_group_f(M.treef, top, ntop, tg, inf, d, k, lo[g], hi[g]);
            if (n >= 0) {
                A[g] = d;
                leaf[g] = node(n).leaf;
                rank[g] = M.leaf_rank[leaf[g]];
            }
        }
    }
    for (int qy = 0; qy < kQH; ++qy) {
        const size_t i = col + static_cast<size_t>(qy) * 16;
        const int64_t tv = T.at(i, k);
        if (!scanned && !in_range(tv)) {
            lstar[i] = -2;
            mind[i] = -1;
            continue;
        }
        const float v = static_cast<float>(tv);
        double best = inf;
        int bl = -1, br = INT32_MAX;
#pragma unroll
        for (int g = 0; g < kColGroups; ++g) {
            const float p = fmaxf(fmaxf(v - hi[g], lo[g] - v), 0.0f);
            const double d = A[g] + static_cast<double>(p) * static_cast<double>(p);
            if (leaf[g] >= 0 && (d < best || (d == best && rank[g] < br))) {
                best = d;
                bl = leaf[g];
                br = rank[g];
            }
        }
        lstar[i] = bl;
        mind[i] = static_cast<int64_t>(best);
    }
}

// The column search in two launches with more threads: the group searches of all columns
// (one thread per column and group), then every quart's L* from its column's group results
// (one thread per quart). Used when the code generator shows that at most one climate value
// varies along a quart column (kVaryCoord) and that coordinate has at most kColGroups groups.
struct GroupRes {
    double a;        // minimal distance over the other coordinates within the group
    int32_t leaf;    // its first leaf in depth-first order; -1: none; -3: the column took the per-quart search
    int32_t rank;    // that leaf's depth-first rank
};

// Group g of quart column c (idx = c * G + g).
MCT_HD inline void biome_group_search(const ModelDev& M, const TreeNodeF* top, int ntop, const Targets& T, int32_t* lstar, int64_t* mind,
                                      GroupRes* out, int G, int idx) {
    const int colidx = idx / G, g = idx - colidx * G, k = kVaryCoord;
    const int b = colidx >> 4, qx = colidx & 3, qz = (colidx >> 2) & 3;
    const size_t col = static_cast<size_t>(b) * kQPer + qz * 4 + qx;
    int64_t t0[7];
    T.get(col, t0);
    bool small = M.treef != nullptr;
    for (int j = 0; j < 7; ++j)
        if (j != k) small = small && t0[j] >= -static_cast<int64_t>(kTreeExact) && t0[j] <= static_cast<int64_t>(kTreeExact);
    if (!small) {
        if (g == 0) biome_search(M, top, ntop, T, lstar, mind, colidx, false, kQH);
        out[idx] = GroupRes{0.0, -3, 0};
        return;
    }
    const double inf = u2d(0x7FF0000000000000ull);
    const float lo = M.col_lo[k][g], hi = M.col_hi[k][g];
    float tg[7];
    for (int j = 0; j < 7; ++j) tg[j] = j == k ? lo : static_cast<float>(t0[j]);
    double d = 0.0;
    int n = tree_first_min_group_f(M.treef, top, ntop, tg, tree_greedy_f(M.treef, top, ntop, tg) + 1.0, d, k, lo, hi);
    if (n < 0) n = tree_first_min_group_f(M.treef, top, ntop, tg, inf, d, k, lo, hi);
    if (n < 0) {
        out[idx] = GroupRes{inf, -1, INT32_MAX};
    } else {
        const int leaf = n < ntop ? top[n].leaf : M.treef[n].leaf;
        out[idx] = GroupRes{d, leaf, M.leaf_rank[leaf]};
    }
}

// L* and its distance of quart idx (biome chunk * kQPer + quart) from its column's G groups.
// tvary: each quart's varying coordinate (biome_target's tvary).
MCT_HD inline void biome_quart_final(const ModelDev& M, const int64_t* tvary, const GroupRes* res, int32_t* lstar, int64_t* mind, int G, int idx) {
    const int b = idx / kQPer, q = idx - b * kQPer;
    const GroupRes* r = res + static_cast<size_t>(b * 16 + (q & 15)) * G;
    // every load of the quart at once (G <= kColGroups; tvary holds every quart's coordinate)
    const int64_t tv = tvary[idx];
    GroupRes rg[kColGroups];
#pragma unroll
    for (int g = 0; g < kColGroups; ++g)
        if (g < G) rg[g] = r[g];
    if (rg[0].leaf == -3) return;
    const int k = kVaryCoord;
    if (tv < -static_cast<int64_t>(kTreeExact) || tv > static_cast<int64_t>(kTreeExact)) {
        lstar[idx] = -2;
        mind[idx] = -1;
        return;
    }
    const float v = static_cast<float>(tv);
    double best = u2d(0x7FF0000000000000ull);
    int bl = -1, br = INT32_MAX;
#pragma unroll
    for (int g = 0; g < kColGroups; ++g) {
        if (g >= G) break;
        const float p = fmaxf(fmaxf(v - M.col_hi[k][g], M.col_lo[k][g] - v), 0.0f);
        const double d = rg[g].a + static_cast<double>(p) * static_cast<double>(p);
        if (rg[g].leaf >= 0 && (d < best || (d == best && rg[g].rank < br))) {
            best = d;
            bl = rg[g].leaf;
            br = rg[g].rank;
        }
    }
    lstar[idx] = bl;
    mind[idx] = static_cast<int64_t>(best);
}

// Chain order of doCreateBiomes inside a chunk: 4-quart sections, then x, y, z.
MCT_HD MCT_FI int chain_quart(int p) {
    const int z = p & 3, y = (p >> 2) & 3, x = (p >> 4) & 3, sy = p >> 6;
    return ((sy * 4 + y) * 4 + z) * 4 + x;
}

// ---- biome-only queries -------------------------------------------------------------------
// What the TERRAIN step stores for block (x, y, z): the zoom picks a quart (y clamped to the
// chunk's quart rows), and the answer is that quart's biome as the chunk's chain of R-tree
// searches leaves it: result(p) = result(p-1) when it is also at the minimal distance at p,
// else L*_p, with result(-1) = none. The query walks the chain back from p to the nearest
// position whose minimum is unique (or the chain start), where the result cannot depend on
// the history, then replays the rule forward.
struct QuartSearch {
    float t[7];
    int leaf;        // L*
    double d;        // its distance
    bool unique;     // no other leaf at distance d
};

// Position p of the chain order of chunk (cx, cz): climate target, L*, uniqueness.
MCT_HD inline bool quart_search(const SeedDev& S, const ModelDev& M, int cx, int cz, int p, QuartSearch& out) {
    const int qi = chain_quart(p);
    const int qx = qi & 3, qz = (qi >> 2) & 3, qy = qi >> 4;
    float cv[mctgen::kNumClimateColumns > 0 ? mctgen::kNumClimateColumns : 1];
    mctgen::eval_climate_columns(S.oct, cx * 16 + qx * 4, cz * 16 + qz * 4, cv);
    float clim[6];
    mctgen::eval_climate(S.oct, cx * 16 + qx * 4, kMinY + qy * 4, cz * 16 + qz * 4, cv, clim);
    for (int k = 0; k < 7; ++k) {
        const int64_t v = k < 6 ? quantize(clim[k]) : 0;
        if (!M.treef || v < -static_cast<int64_t>(kTreeExact) || v > static_cast<int64_t>(kTreeExact)) return false;
        out.t[k] = static_cast<float>(v);
    }
    const double inf = u2d(0x7FF0000000000000ull);
    out.leaf = tree_first_min_f(M.treef, nullptr, 0, out.t, inf, out.d);
    double d2 = 0.0;
    const int other = tree_first_min_f(M.treef, nullptr, 0, out.t, out.d + 1.0, d2, M.leaf_node[out.leaf]);
    out.unique = other < 0;
    return true;
}

constexpr int kQueryWalk = 16;

// Biome stored for block (x, y, z) under seed table S; -1 when the float search cannot be used
// (the caller then generates the chunk's biomes the ordinary way).
MCT_HD inline int biome_query(const SeedDev& S, const ModelDev& M, int32_t x, int32_t y, int32_t z) {
    int32_t qx, qy, qz;
    zoom_pick(S.zoom_seed, x, y, z, qx, qy, qz);
    const int qmin = kMinY >> 2;
    qy = clampi(qy, qmin, qmin + kQH - 1) - qmin;
    const int cx = qx >> 2, cz = qz >> 2;
    const int p = ((qy >> 2) << 6) | ((qx & 3) << 4) | ((qy & 3) << 2) | (qz & 3);   // chain position
    QuartSearch w[kQueryWalk];
    int n = 0, anchor = -1;
    for (int k = p; k >= 0 && n < kQueryWalk; --k, ++n) {
        if (!quart_search(S, M, cx, cz, k, w[n])) return -1;
        if (w[n].unique || k == 0) { anchor = n; break; }
    }
    int r;
    if (anchor >= 0) {
        r = w[anchor].leaf;
        for (int j = anchor - 1; j >= 0; --j) {
            const double dp = box_distance_f(M.treef[M.leaf_node[r]], w[j].t);
            if (!(dp == w[j].d)) r = w[j].leaf;
        }
    } else {   // no anchor within the walk: replay the chain from its start
        r = -1;
        QuartSearch q;
        for (int k = 0; k <= p; ++k) {
            if (!quart_search(S, M, cx, cz, k, q)) return -1;
            r = (r >= 0 && box_distance_f(M.treef[M.leaf_node[r]], q.t) == q.d) ? r : q.leaf;
        }
    }
    return M.leaf_biome[r];
}

// tie[p]: the previous chain position's L* differs from L* at p and is also at minimal
// distance at p. While the chain is in sync (previous result == previous L*), the result at p
// differs from L* exactly where tie[p] is set.
MCT_HD MCT_FI void biome_tie(const ModelDev& M, const Targets& T, const int32_t* lstar, const int64_t* mind, uint8_t* tie, int idx) {
    const int b = idx / kQPer, p = idx % kQPer;
    if (p == 0) { tie[idx] = 0; return; }
    const size_t i = static_cast<size_t>(b) * kQPer + chain_quart(p), ip = static_cast<size_t>(b) * kQPer + chain_quart(p - 1);
    const int prev = lstar[ip];
    bool t = false;
    if (prev != lstar[i]) {
        int64_t tg[7];
        T.get(i, tg);
        t = box_distance(M.tree[M.leaf_node[prev]], tg) == mind[i];
    }
    tie[idx] = t ? 1 : 0;
}

// The last-result rule along the chain of one biome chunk (sequential form): the result is
// the previous result if it reaches the minimal distance, else L*.
MCT_HD inline void biome_chain_seq(const ModelDev& M, const Targets& T, const int32_t* lstar, const int64_t* mind, uint8_t* biomes, int b) {
    const size_t base = static_cast<size_t>(b) * kQPer;
    int prev = -1;
    for (int p = 0; p < kQPer; ++p) {
        const size_t i = base + chain_quart(p);
        int cur = lstar[i];
        int64_t tg[7];
        T.get(i, tg);
        if (prev >= 0 && box_distance(M.tree[M.leaf_node[prev]], tg) == mind[i]) cur = prev;
        biomes[i] = M.leaf_biome[cur];
        prev = cur;
    }
}

// ---- density ---------------------------------------------------------------------------
MCT_HD MCT_FI void column_key(const KeyCodec& K, const ChunkRef* ch, uint64_t* keys, int idx) {
    const int c = idx / kCols, k = idx % kCols;
    keys[idx] = make_key(K, ch[c].seed, 0, ch[c].x * 4 + k % 5, ch[c].z * 4 + k / 5);
}

MCT_HD MCT_FI void column_values(const KeyCodec& K, const SeedDev* seeds, const uint64_t* ucols, float* colv, int u) {
    int32_t s, qx, qz;
    int f;
    split_key(K, ucols[u], s, f, qx, qz);
    MCT_OPS(0, 0, 8, 1);   // the key read and split, x4 coordinates
    mctgen::eval_columns(seeds[s].oct, qx * 4, qz * 4, colv + static_cast<size_t>(u) * (mctgen::kNumColumns > 0 ? mctgen::kNumColumns : 1));
}

// x / z records of corner-pass octave sample site `site` of distinct column u.
MCT_HD MCT_FI mct::AxisXZ corner_site_axes(const KeyCodec& K, const SeedDev* seeds, const uint64_t* ucols, int u, int site) {
    int32_t s, qx, qz;
    int f;
    split_key(K, ucols[u], s, f, qx, qz);
    return mctgen::corner_axes(seeds[s].oct, qx * 4, qz * 4, site);
}

// Channel inputs at one cell corner (column u, corner row): corners[(u * kCh + ch) * kRows + row].
// axz: the column's site records (corner_site_axes for every site).
MCT_HD MCT_FI void corner_values(const KeyCodec& K, const SeedDev* seeds, const uint64_t* ucols, const float* colv, float* corners, int u, int row,
                                 const MCT_LDS mct::AxisXZ* axz) {
    int32_t s, qx, qz;
    int f;
    split_key(K, ucols[u], s, f, qx, qz);
    MCT_OPS(0, 0, 9, 1);   // the key read and split, x4 coordinates and the corner's y
    float out[kCh];
    mctgen::eval_corner(seeds[s].oct, qx * 4, kMinY + row * 8, qz * 4, colv + static_cast<size_t>(u) * (mctgen::kNumColumns > 0 ? mctgen::kNumColumns : 1), axz,
                        seeds[s].ytab + row, out);
    for (int ch = 0; ch < kCh; ++ch) corners[(static_cast<size_t>(u) * kCh + ch) * kRows + row] = out[ch];
}

// Channel values of a block by the cell fill (z, then x, then the y steps).
MCT_HD MCT_FI void interp_channels(const float* corners, const int* cmap, int x, int yy, int z, int first, int count, float* out) {
    const int cx = x >> 2, xi = x & 3, cz = z >> 2, zi = z & 3, cy = yy >> 3, yi = yy & 7;
    const size_t u00 = cmap[cz * 5 + cx], u10 = cmap[cz * 5 + cx + 1], u01 = cmap[(cz + 1) * 5 + cx], u11 = cmap[(cz + 1) * 5 + cx + 1];
    for (int k = 0; k < count; ++k) {
        const int ch = first + k;
        const float* a = corners + (u00 * kCh + ch) * kRows;
        const float* b = corners + (u10 * kCh + ch) * kRows;
        const float* c = corners + (u01 * kCh + ch) * kRows;
        const float* d = corners + (u11 * kCh + ch) * kRows;
        out[k] = cell_fill(a[cy], b[cy], a[cy + 1], b[cy + 1], c[cy], d[cy], c[cy + 1], d[cy + 1], xi, yi, zi, 0.25f, 0.125f);
    }
}

// ---- surface-level grid (aquifer) and preliminary surface corners ------------------------
// Per chunk: 27 x 19 surface-level quarts around it; inside NoiseBasedAquifer's constructor
// prefill window (quarts 4c-4 .. 4c+6) the volume path is used, elsewhere the point path;
// plus the 4 chunk corners of chunk_surface_level (volume path).
MCT_HD MCT_FI void grid_key(const KeyCodec& K, const ChunkRef* ch, uint64_t* keys, int idx) {
    const int c = idx / kGridKeys, e = idx % kGridKeys;
    const int cx = ch[c].x, cz = ch[c].z;
    if (e < kGridCells) {
        const int qx = 4 * cx - 16 + e % kGridX, qz = 4 * cz - 8 + e / kGridX;
        const bool vol = qx >= 4 * cx - 4 && qx <= 4 * cx + 6 && qz >= 4 * cz - 4 && qz <= 4 * cz + 6;
        keys[idx] = make_key(K, ch[c].seed, vol ? 1 : 0, qx, qz);
    } else {
        const int k = e - kGridCells;
        keys[idx] = make_key(K, ch[c].seed, 1, 4 * cx + 4 * (k & 1), 4 * cz + 4 * (k >> 1));
    }
}

MCT_HD MCT_FI void fts_value(const KeyCodec& K, const SeedDev* seeds, const uint64_t* ukeys, float* vals, int u) {
    int32_t s, qx, qz;
    int vol;
    split_key(K, ukeys[u], s, vol, qx, qz);
    MCT_OPS(0, 0, 8, 1);
    const Octave* O = seeds[s].oct;
    vals[u] = vol ? mctgen::fts_volume(O, qx * 4, qz * 4) : mctgen::fts_point(O, qx * 4, qz * 4);
}

MCT_HD MCT_FI void grid_scatter(const int* map, const float* vals, int32_t* grid, float* prelim, int idx) {
    const int c = idx / kGridKeys, e = idx % kGridKeys;
    const float v = vals[map[idx]];
    if (e < kGridCells) grid[static_cast<size_t>(c) * kGridCells + e] = mfloor(static_cast<double>(v));
    else prelim[static_cast<size_t>(c) * 4 + (e - kGridCells)] = v;
}

// ---- aquifer (NoiseBasedAquifer) ----------------------------------------------------------
struct Fluid {
    int level, type;
};
MCT_HD MCT_FI int fluid_at(Fluid f, int y) { return y < f.level ? f.type : mctgen::kAir; }
MCT_HD MCT_FI Fluid global_fluid(int y) {
    const int lava_level = kSea < -54 ? kSea : -54;
    return y < lava_level ? Fluid{-54, mctgen::kLava} : Fluid{kSea, mctgen::kDefaultFluid};
}
MCT_HD MCT_FI int grid_yf(int b) { return fdiv(b, 12); }

// Skip height: blocks above it take the global fluid directly (max surface level over the
// prefill window, +8, up one grid cell).
MCT_HD MCT_FI void aquifer_skip(const int32_t* grid, int* skip, int c) {
    int m = INT32_MIN;
    for (int gz = 4; gz <= 14; ++gz)
        for (int gx = 12; gx <= 22; ++gx) m = jmaxi(m, grid[static_cast<size_t>(c) * kGridCells + gz * kGridX + gx]);
    const int adjusted = m + 8;
    const int skip_gy = grid_yf(adjusted + 12) + 1;
    skip[c] = skip_gy * 12 + 11 - 1;
}

MCT_HD MCT_FI int surface_at(const int32_t* g, int cx, int cz, int x, int z) {
    return g[((x >> 2) - (4 * cx - 16)) + ((z >> 2) - (4 * cz - 8)) * kGridX];
}

// Location (aquifer random) and computeFluid status of one aquifer cell of a chunk.
MCT_HD inline void aquifer_cell(const SeedDev* seeds, const ChunkRef* ch, const int32_t* grid, int32_t* cell_loc, int32_t* cell_fluid, int idx) {
    const int c = idx / kAqCells, i = idx % kAqCells;
    const int cx = ch[c].x, cz = ch[c].z;
    const SeedDev& S = seeds[ch[c].seed];
    const int lx = i % 3, lz = (i / 3) % 3, ly = i / 9;
    const int gx = cx - 1 + lx, gz = cz - 1 + lz, gy = kAqMinGY + ly;
    Xoro r = at(S.aquifer, gx, gy, gz);
    const int x = (gx << 4) + next_int(r, 10);
    const int y = gy * 12 + next_int(r, 9);
    const int z = (gz << 4) + next_int(r, 10);
    const int32_t* g = grid + static_cast<size_t>(c) * kGridCells;
    const Fluid global = global_fluid(y);
    Fluid res = global;
    bool done = false;
    int lowest = INT32_MAX;
    const int top = y + 12, bottom = y - 12;
    bool center_under = false;
    const int off[13][2] = {{0, 0}, {-2, -1}, {-1, -1}, {0, -1}, {1, -1}, {-3, 0}, {-2, 0}, {-1, 0}, {1, 0}, {-2, 1}, {-1, 1}, {0, 1}, {1, 1}};
    for (int k = 0; k < 13; ++k) {
        const int sx = x + (off[k][0] << 4), sz = z + (off[k][1] << 4);
        const int level = surface_at(g, cx, cz, sx, sz);
        const int adjusted = level + 8;
        const bool start = k == 0;
        if (start && bottom > adjusted) { res = global; done = true; break; }
        const bool pokes = top > adjusted;
        if (pokes || start) {
            const Fluid at_surface = global_fluid(adjusted);
            if (fluid_at(at_surface, adjusted) != mctgen::kAir) {
                if (start) center_under = true;
                if (pokes) { res = at_surface; done = true; break; }
            }
        }
        lowest = jmini(lowest, level);
    }
    if (!done) {
        const Octave* O = S.oct;
        double partially, fully;
        if (mctgen::eval_point(O, x, y, z, 4) > 0.0f) {
            partially = -1.0;
            fully = -1.0;
        } else {
            const int below = lowest + 8 - y;
            double factor = 0.0;
            if (center_under) {
                const double f = (static_cast<double>(below) - 0.0) / (64.0 - 0.0);
                factor = f < 0.0 ? 1.0 : (f > 1.0 ? 0.0 : 1.0 + f * (0.0 - 1.0));
            }
            const double noise = clampd(static_cast<double>(mctgen::eval_point(O, x, y, z, 1)), -1.0, 1.0);
            const double fully_t = -0.3 + ((factor - 1.0) / (0.0 - 1.0)) * (0.8 - -0.3);
            const double partial_t = -0.8 + ((factor - 1.0) / (0.0 - 1.0)) * (0.4 - -0.8);
            partially = noise - partial_t;
            fully = noise - fully_t;
        }
        int level;
        if (fully > 0.0) {
            level = global.level;
        } else if (partially > 0.0) {
            const int fcx = fdiv(x, 16), fcy = fdiv(y, 40), fcz = fdiv(z, 16);
            const int middle = fcy * 40 + 20;
            const double spread = static_cast<double>(mctgen::eval_point(O, fcx, fcy, fcz, 2) * 10.0f);
            const int q = mfloor(spread / 3.0) * 3;
            level = jmini(lowest, middle + q);
        } else {
            level = kWayBelow;
        }
        int type = global.type;
        if (level <= -10 && level != kWayBelow && global.type != mctgen::kLava) {
            const int tcx = fdiv(x, 64), tcy = fdiv(y, 40), tcz = fdiv(z, 64);
            const double lava = static_cast<double>(mctgen::eval_point(O, tcx, tcy, tcz, 3));
            if (fabs(lava) > 0.3) type = mctgen::kLava;
        }
        res = Fluid{level, type};
    }
    cell_loc[static_cast<size_t>(idx) * 4 + 0] = x;
    cell_loc[static_cast<size_t>(idx) * 4 + 1] = y;
    cell_loc[static_cast<size_t>(idx) * 4 + 2] = z;
    cell_loc[static_cast<size_t>(idx) * 4 + 3] = 0;
    cell_fluid[static_cast<size_t>(idx) * 2 + 0] = res.level;
    cell_fluid[static_cast<size_t>(idx) * 2 + 1] = res.type;
}

// x / d for the pressures' divisions, whose x are half-integers of magnitude below 2^18
// (fluid levels, y and 0.5 offsets; levels lie in [kWayBelow, kHeight]) and whose d are 25,
// 1.5, 2.5, 3 and 10. On the device div_known, equal to the IEEE division on every
// half-integer of magnitude up to 2^19 for these d (checked exhaustively on the CPU).
MCT_HD MCT_FI double aq_div(double x, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    return div_known(x, d, 1.0 / d);
#else
    return x / d;
#endif
}

MCT_HD MCT_FI double aquifer_pressure(const Octave* O, int x, int y, int z, double& barrier, bool& barrier_set, Fluid a, Fluid b) {
    const int t1 = fluid_at(a, y), t2 = fluid_at(b, y);
    MCT_OPS(0, 0, 6, 0);   // two fluid picks, the lava / water pair test
    if ((t1 == mctgen::kLava && t2 == mctgen::kWater) || (t1 == mctgen::kWater && t2 == mctgen::kLava)) return 2.0;
    const int diff = a.level - b.level < 0 ? b.level - a.level : a.level - b.level;
    MCT_OPS(0, 0, 4, 0);
    if (diff == 0) return 0.0;
    MCT_OPS(14, 0, 1, 0);   // average, above, base, towards, the gradient branch and the range test
    const double average = 0.5 * static_cast<double>(a.level + b.level);
    const double above = static_cast<double>(y) + 0.5 - average;
    const double base = static_cast<double>(diff) / 2.0;
    const double towards = base - fabs(above);
    double gradient;
    if (above > 0.0) {
        const double center = 0.0 + towards;
        gradient = center > 0.0 ? aq_div(center, 1.5) : aq_div(center, 2.5);
    } else {
        const double center = 3.0 + towards;
        gradient = center > 0.0 ? aq_div(center, 3.0) : aq_div(center, 10.0);
    }
    double noise;
    if (!(gradient < -2.0) && !(gradient > 2.0)) {
        if (!barrier_set) {
            barrier = static_cast<double>(mctgen::eval_point(O, x, y, z, 0));
            barrier_set = true;
        }
        noise = barrier;
    } else {
        noise = 0.0;
    }
    MCT_OPS(2, 0, 0, 0);
    return 2.0 * (noise + gradient);
}

// The aquifer cells of a chunk as computeSubstance reads them: cell i's location (x, y, z)
// and its fluid status. Cell index i = ((gy - kAqMinGY) * 3 + (gz - (cz - 1))) * 3 + (gx - (cx - 1)).
struct AqCellsGlobal {
    const int32_t* loc;   // 4 ints per cell (x, y, z, 0)
    const int32_t* fl;    // 2 ints per cell (level, fluid)
    MCT_HD MCT_FI void at(int i, int& x, int& y, int& z) const { x = loc[i * 4]; y = loc[i * 4 + 1]; z = loc[i * 4 + 2]; }
    MCT_HD MCT_FI Fluid fluid(int i) const { return Fluid{fl[i * 2], fl[i * 2 + 1]}; }
};

// dx^2 + dy^2 + dz^2 of offsets within +-2^23: the products are exact in 24-bit signed
// multiplies, which issue at the full VALU rate (the 32-bit multiply is quarter rate).
MCT_HD MCT_FI int sq3_small(int dx, int dy, int dz) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __mul24(dx, dx) + __mul24(dy, dy) + __mul24(dz, dz);
#else
    return dx * dx + dy * dy + dz * dz;
#endif
}

// The three nearest aquifer cells of a block (computeSubstance's 2 x 3 x 2 scan, ties to the
// later cell of the scan). A cell of the scan lies within
// [16 * gx, 16 * gx + 9] x [12 * gy, 12 * gy + 8] x [16 * gz, 16 * gz + 9] (aquifer_cell) and
// gx, gy, gz are within one cell of the block's, so every offset is at most 22 in size.
struct AqNearest {
    int i1, i2, i3, d1, d2, d3;
};
template <class Cells>
MCT_HD MCT_FI AqNearest aquifer_nearest(const Cells& C, int cx, int cz, int x, int y, int z) {
    const int xa = (x - 5) >> 4, ya = grid_yf(y + 1), za = (z - 5) >> 4;
    MCT_OPS(0, 0, 7, 0);
    int d1 = INT32_MAX, d2 = INT32_MAX, d3 = INT32_MAX, d4 = INT32_MAX;
    int i1 = 0, i2 = 0, i3 = 0;
    for (int ox = 0; ox <= 1; ++ox)
        for (int oy = -1; oy <= 1; ++oy)
            for (int oz = 0; oz <= 1; ++oz) {
                const int gx = xa + ox, gy = ya + oy, gz = za + oz;
                // cell index (9), offsets and squared distance (8), the insertion compares (4)
                MCT_OPS(0, 0, 21, 3);
                const int index = ((gy - kAqMinGY) * 3 + (gz - (cz - 1))) * 3 + (gx - (cx - 1));
                int lx, ly, lz;
                C.at(index, lx, ly, lz);
                const int dx = lx - x, dy = ly - y, dz = lz - z;
                const int nd = sq3_small(dx, dy, dz);
                if (d1 >= nd) { i3 = i2; i2 = i1; i1 = index; d4 = d3; d3 = d2; d2 = d1; d1 = nd; }
                else if (d2 >= nd) { i3 = i2; i2 = index; d4 = d3; d3 = d2; d2 = nd; }
                else if (d3 >= nd) { i3 = index; d4 = d3; d3 = nd; }
                else if (d4 >= nd) { d4 = nd; }
            }
    return AqNearest{i1, i2, i3, d1, d2, d3};
}

// NoiseBasedAquifer.computeSubstance up to the barrier: the result, or kAqPending where the
// barrier / pressure part decides (aquifer_back). -1 means solid (the default block).
// sim12 = 1 - (d2 - d1) / 25 is <= 0 exactly when d2 - d1 >= 25 (the quotient of integers
// 0..24 by 25 rounds below 1, of 25 and up to 1 or more), so the front compares integers.
constexpr int kAqPending = -2;
template <class Cells>
MCT_HD MCT_FI int aquifer_front(const Cells& C, int cx, int cz, int skip_above, int x, int y, int z, double density, AqNearest& nr) {
    MCT_OPS(1, 0, 0, 0);
    if (density > 0.0) return -1;
    const Fluid global = global_fluid(y);
    MCT_OPS(0, 0, 4, 0);   // the global fluid pick, the skip test
    if (y > skip_above) return fluid_at(global, y);
    if (fluid_at(global, y) == mctgen::kLava) return mctgen::kLava;
    nr = aquifer_nearest(C, cx, cz, x, y, z);
    const int fluid = fluid_at(C.fluid(nr.i1), y);
    MCT_OPS(0, 0, 4, 2);   // the nearest cell's fluid, the d2 - d1 test
    if (nr.d2 - nr.d1 >= 25) return fluid;
    MCT_OPS(0, 0, 4, 0);
    if (fluid == mctgen::kWater && fluid_at(global_fluid(y - 1), y - 1) == mctgen::kLava) return fluid;
    return kAqPending;
}

// The rest of computeSubstance for a block the front left pending (sim12 > 0, not water over
// the lava level): the barrier-weighted pressures of the cell pairs 1-2, 1-3, 2-3.
template <class Cells>
MCT_HD inline int aquifer_back(const Octave* O, const Cells& C, int x, int y, int z, double density, const AqNearest& nr) {
    const Fluid s1 = C.fluid(nr.i1);
    const int fluid = fluid_at(s1, y);
    const double sim12 = 1.0 - aq_div(static_cast<double>(nr.d2 - nr.d1), 25.0);
    MCT_OPS(5, 0, 2, 4);   // sim12, b12 and its test; the fluids of cells 1 and 2
    double barrier = 0.0;
    bool barrier_set = false;
    const Fluid s2 = C.fluid(nr.i2);
    const double b12 = sim12 * aquifer_pressure(O, x, y, z, barrier, barrier_set, s1, s2);
    if (density + b12 > 0.0) return -1;
    const Fluid s3 = C.fluid(nr.i3);
    const double sim13 = 1.0 - aq_div(static_cast<double>(nr.d3 - nr.d1), 25.0);
    MCT_OPS(4, 0, 1, 2);   // sim13 and its test; the fluid of cell 3
    if (sim13 > 0.0) {
        MCT_OPS(4, 0, 0, 0);
        const double b13 = sim12 * sim13 * aquifer_pressure(O, x, y, z, barrier, barrier_set, s1, s3);
        if (density + b13 > 0.0) return -1;
    }
    const double sim23 = 1.0 - aq_div(static_cast<double>(nr.d3 - nr.d2), 25.0);
    MCT_OPS(4, 0, 1, 0);
    if (sim23 > 0.0) {
        MCT_OPS(4, 0, 0, 0);
        const double b23 = sim12 * sim23 * aquifer_pressure(O, x, y, z, barrier, barrier_set, s2, s3);
        if (density + b23 > 0.0) return -1;
    }
    return fluid;
}

// NoiseBasedAquifer.computeSubstance; -1 means solid (the default block).
MCT_HD inline int aquifer_substance(const Octave* O, int cx, int cz, const int32_t* loc, const int32_t* fl, int skip_above, int x, int y, int z,
                                    double density) {
    const AqCellsGlobal C{loc, fl};
    AqNearest nr;
    const int s = aquifer_front(C, cx, cz, skip_above, x, y, z, density, nr);
    return s == kAqPending ? aquifer_back(O, C, x, y, z, density, nr) : s;
}

// ---- fill: one block -------------------------------------------------------------------
struct FillArgs {
    const SeedDev* seeds;
    const ChunkRef* chunks;
    const float* corners;
    const int* colmap;          // n x 25 corner-column indices
    const int32_t* cell_loc;    // n x kAqCells x 4
    const int32_t* cell_fluid;  // n x kAqCells x 2
    const int* skip;
    uint8_t* blocks;
    float* density;             // core configuration: final density per block (nullable)
    int* h0;                    // n x 256 initial heights (column x * 16 + z), raised here; nullable
    uint32_t* clsw;             // n x 24 x 256 block-class words (see SurfArgs::clsw); nullable
};

// The 8 blocks of one block column inside one 4x8 cell (t = (cy * 16 + z) * 16 + x): the z
// and x lerps of every channel once, then y as the running sum of the step, which is the
// same sequence of float additions the per-block form repeats from the cell's bottom.
constexpr int kCellsY = kHeight / 8;

// Channel values at the bottom block of cell row cy of block column (x, z) and their per-block
// y step: the z lerps, then the x lerps, of the cell's 8 corners.
MCT_HD MCT_FI void cell_channels(const FillArgs& A, int c, int x, int z, int cy, float* val, float* step) {
    const int* cmap = A.colmap + static_cast<size_t>(c) * kCols;
    const int cx = x >> 2, xi = x & 3, cz = z >> 2, zi = z & 3;
    const size_t u00 = cmap[cz * 5 + cx], u10 = cmap[cz * 5 + cx + 1], u01 = cmap[(cz + 1) * 5 + cx], u11 = cmap[(cz + 1) * 5 + cx + 1];
    const float az = static_cast<float>(zi) * 0.25f, ax = static_cast<float>(xi) * 0.25f;
    // per column of a cell: the 4 corner columns, the lerp factors; per channel 8 corner values,
    // 6 lerps and the y step
    MCT_OPS(0, 4 + 20 * mctgen::kNumChannels, 12, 4 + 8 * mctgen::kNumChannels);
    // unrolled, so that all channels' corner loads are in flight together (a rolled loop
    // waits for each channel's loads before it issues the next channel's)
#pragma unroll
    for (int k = 0; k < mctgen::kNumChannels; ++k) {
        const float* a = A.corners + (u00 * kCh + k) * kRows + cy;
        const float* b = A.corners + (u10 * kCh + k) * kRows + cy;
        const float* cc = A.corners + (u01 * kCh + k) * kRows + cy;
        const float* d = A.corners + (u11 * kCh + k) * kRows + cy;
        const float v00 = lerpf(az, a[0], cc[0]), v01 = lerpf(az, a[1], cc[1]);
        const float v10 = lerpf(az, b[0], d[0]), v11 = lerpf(az, b[1], d[1]);
        const float lo = lerpf(ax, v00, v10), hi = lerpf(ax, v01, v11);
        step[k] = (hi - lo) * 0.125f;
        val[k] = lo + step[k] * 0.0f;
    }
}

MCT_HD MCT_FI void fill_cell(const FillArgs& A, int c, int t, float* density_out) {
    const int x = t & 15, z = (t >> 4) & 15, cy = t >> 8;
    constexpr int NC = mctgen::kNumChannels > 0 ? mctgen::kNumChannels : 1;
    float val[NC], step[NC];
    cell_channels(A, c, x, z, cy, val, step);
    const ChunkRef cr = A.chunks[c];
    const Octave* O = A.seeds[cr.seed].oct;
    const int bx = cr.x * 16 + x, bz = cr.z * 16 + z;
    const int32_t* loc = A.cell_loc + static_cast<size_t>(c) * kAqCells * 4;
    const int32_t* fl = A.cell_fluid + static_cast<size_t>(c) * kAqCells * 2;
    const int skip = A.skip[c];
    uint8_t* out = A.blocks + static_cast<size_t>(c) * kPer + static_cast<size_t>(cy * 8 * 256 + z * 16 + x);
    int top = -1;
    for (int yi = 0; yi < 8; ++yi) {
        const int by = kMinY + cy * 8 + yi;
        const float d = mctgen::eval_block(O, bx, by, bz, val);
        if (density_out) density_out[(cy * 8 + yi) * 256 + z * 16 + x] = d;
        const int s = aquifer_substance(O, cr.x, cr.z, loc, fl, skip, bx, by, bz, static_cast<double>(d));
        const int state = s < 0 ? mctgen::kDefaultBlock : s;
        out[yi * 256] = static_cast<uint8_t>(state);
        if (!((mctgen::kAirMask >> state) & 1u)) top = yi;
        MCT_OPS(0, mctgen::kNumChannels, 4, 0);   // the block's state and air test, the channel y steps
        for (int k = 0; k < mctgen::kNumChannels; ++k) val[k] += step[k];
    }
#if defined(__HIP_DEVICE_COMPILE__)
    // WORLD_SURFACE_WG after the fill: the highest non-air block + 1 of the column, the max over
    // its cells (the heights start at minY; a plain read skips most of the atomics)
    if (top >= 0 && A.h0) {
        int* h = A.h0 + static_cast<size_t>(c) * 256 + x * 16 + z;
        const int v = kMinY + cy * 8 + top + 1;
        if (v > *reinterpret_cast<volatile int*>(h)) atomicMax(h, v);
    }
#endif
}

// Block ids of the core configuration's fill (no aquifers, no material rules).
enum CoreBlock : uint8_t { kCoreAir = 0, kCoreStone = 1, kCoreWater = 2, kCoreLava = 3 };

// Vanilla doFill substance with aquifers disabled: the default block where density > 0, else
// the global fluid picker (lava below min(-54, sea level) and y < -54, water below sea level).
// Reference: cpu/include/mccuda/volume_eval.h:76-85.
MCT_HD MCT_FI uint8_t core_fill(float density, int y) {
    if (static_cast<double>(density) > 0.0) return kCoreStone;
    const int lava_level = kSea < -54 ? kSea : -54;
    if (y < lava_level) return y < -54 ? kCoreLava : kCoreAir;
    return y < kSea ? kCoreWater : kCoreAir;
}

// Core configuration, one block column of one cell: final density (A.density) and core fill
// (A.blocks), either may be null. Same channel interpolation and density as fill_cell.
MCT_HD MCT_FI void fill_cell_core(const FillArgs& A, int c, int t) {
    const int x = t & 15, z = (t >> 4) & 15, cy = t >> 8;
    constexpr int NC = mctgen::kNumChannels > 0 ? mctgen::kNumChannels : 1;
    float val[NC], step[NC];
    cell_channels(A, c, x, z, cy, val, step);
    const ChunkRef cr = A.chunks[c];
    const Octave* O = A.seeds[cr.seed].oct;
    const int bx = cr.x * 16 + x, bz = cr.z * 16 + z;
    const size_t at = static_cast<size_t>(c) * kPer + static_cast<size_t>(cy * 8 * 256 + z * 16 + x);
    for (int yi = 0; yi < 8; ++yi) {
        const int by = kMinY + cy * 8 + yi;
        const float d = mctgen::eval_block(O, bx, by, bz, val);
        if (A.density) A.density[at + static_cast<size_t>(yi) * 256] = d;
        if (A.blocks) A.blocks[at + static_cast<size_t>(yi) * 256] = core_fill(d, by);
        for (int k = 0; k < mctgen::kNumChannels; ++k) val[k] += step[k];
    }
}

// Initial WORLD_SURFACE_WG height of a column (highest non-air block + 1, or minY).
MCT_HD MCT_FI int column_height(const ModelDev& M, const uint8_t* blocks, int x, int z) {
    for (int yy = kHeight - 1; yy >= 0; --yy)
        if (!((mctgen::kAirMask >> blocks[(yy * 16 + z) * 16 + x]) & 1u)) return kMinY + yy + 1;
    return kMinY;
}

// ---- legacy simplex noise (biome temperature) -------------------------------------------
MCT_HD MCT_FI double simplex_corner(int g, double x, double y) {
    double t0 = 0.5 - x * x - y * y - 0.0 * 0.0;
    if (t0 < 0.0) return 0.0;
    t0 *= t0;
    double dot = static_cast<double>(gcomp(kGX, g)) * x + static_cast<double>(gcomp(kGY, g)) * y;
    dot = dot + static_cast<double>(gcomp(kGZ, g)) * 0.0;
    return t0 * t0 * dot;
}

MCT_HD inline float simplex2(const uint8_t* perm, double px, double py) {
    const double s3 = 1.7320508075688772;   // Math.sqrt(3.0)
    const double F2 = 0.5 * (s3 - 1.0), G2 = (3.0 - s3) / 6.0;
    const double xin = px + 0.0, yin = py + 0.0;
    const double s = (xin + yin) * F2;
    const int32_t i = mfloor(xin + s), j = mfloor(yin + s);
    const double t = static_cast<double>(iadd(i, j)) * G2;
    const double x0 = xin - (static_cast<double>(i) - t), y0 = yin - (static_cast<double>(j) - t);
    int i1, j1;
    if (x0 > y0) { i1 = 1; j1 = 0; } else { i1 = 0; j1 = 1; }
    const double x1 = x0 - static_cast<double>(i1) + G2, y1 = y0 - static_cast<double>(j1) + G2;
    const double x2 = x0 - 1.0 + 2.0 * G2, y2 = y0 - 1.0 + 2.0 * G2;
    const int ii = i & 0xFF, jj = j & 0xFF;
    const int g0 = perm[(ii + perm[jj]) & 0xFF] % 12;
    const int g1 = perm[(ii + i1 + perm[(jj + j1) & 0xFF]) & 0xFF] % 12;
    const int g2 = perm[(ii + 1 + perm[(jj + 1) & 0xFF]) & 0xFF] % 12;
    return static_cast<float>(70.0 * (simplex_corner(g0, x0, y0) + simplex_corner(g1, x1, y1) + simplex_corner(g2, x2, y2)));
}

// Biome.getHeightAdjustedTemperature, split into what depends on the column only (the
// frozen-modifier patch test and the snow-line noise, kept in the column record) and the rest.
// frozen patch: large (three frozen simplex layers x 7) + edge < 0.3 and small < 0.8.
MCT_HD inline bool frozen_patch(const ModelDev& M, int x, int z) {
    const double fr[3] = {1.0, 0.5, 0.25};
    const float amp[3] = {0.14285715f, 0.2857143f, 0.5714286f};
    float fv = 0.0f;
    for (int k = 0; k < 3; ++k) fv += amp[k] * simplex2(M.simplex[1 + k], static_cast<double>(x) * 0.05 * fr[k], static_cast<double>(z) * 0.05 * fr[k]);
    const double large = static_cast<double>(fv * 7.0f);
    const double edge = static_cast<double>(simplex2(M.simplex[4], static_cast<double>(x) * 0.2, static_cast<double>(z) * 0.2));
    if (!(large + edge < 0.3)) return false;
    const double small = static_cast<double>(simplex2(M.simplex[4], static_cast<double>(x) * 0.09, static_cast<double>(z) * 0.09));
    return small < 0.8;
}

// the temperature noise above the snow line: simplex(x / 8f, z / 8f) * 8f
MCT_HD inline float snow_noise(const ModelDev& M, int x, int z) {
    return simplex2(M.simplex[0], static_cast<double>(static_cast<float>(x) / 8.0f), static_cast<double>(static_cast<float>(z) / 8.0f)) * 8.0f;
}

MCT_HD MCT_FI float biome_temperature_col(const ModelDev& M, int b, int y, bool patch, float snow_v) {
    const float adjusted = (M.biome_frozen[b] && patch) ? 0.2f : M.biome_temp[b];
    const int snow = kSea + 17;
    if (y > snow) {
        float t = snow_v + static_cast<float>(y);
        t = t - static_cast<float>(snow);
        t = t * 0.05f;
        t = t / 40.0f;
        return adjusted - t;
    }
    return adjusted;
}

// Column record: the rule tree's 2D noises (slots of surf_column_noises, including surface,
// surface_secondary and clay_bands_offset), the snow-line noise and the frozen-patch flag.
constexpr int kColF = mctgen::kColNoises + 2;
constexpr int kColSnow = mctgen::kColNoises, kColPatch = mctgen::kColNoises + 1;

// The record of one column (idx = chunk * 256 + lx * 16 + lz).
// Biomes of the 4 quart columns the zoom of block column (x, z) of chunk ci picks from, at any
// height (per quart column: bmask; all bits without it).
MCT_HD MCT_FI void column_candidates(const int* nb_index, const uint64_t* bmask, const ChunkRef& cr, int ci, int x, int z, uint64_t* out) {
    out[0] = out[1] = bmask ? 0 : ~0ull;
    if (!bmask) return;
    const int px = (x - 2) >> 2, pz = (z - 2) >> 2;
    for (int j = 0; j < 4; ++j) {
        const int qx = px + (j >> 1), qz = pz + (j & 1);
        const int bc = nb_index[ci * 9 + ((qz >> 2) - cr.z + 1) * 3 + (qx >> 2) - cr.x + 1];
        const uint64_t* m = bmask + (static_cast<size_t>(bc) * 16 + (qz & 3) * 4 + (qx & 3)) * 2;
        out[0] |= m[0];
        out[1] |= m[1];
    }
}

// The column record: the rule tree's 2D noises (a slot the rules can only read under biomes
// none of the column's zoom candidates has is left 0: it is never read), the snow-line noise
// and the frozen-patch test.
MCT_HD inline void surface_column(const SeedDev* seeds, const ModelDev& M, const ChunkRef* chunks, float* rec, int idx, const int* nb_index = nullptr,
                                  const uint64_t* bmask = nullptr) {
    const int ci = idx >> 8, col = idx & 255;
    const ChunkRef cr = chunks[ci];
    const int x = cr.x * 16 + (col >> 4), z = cr.z * 16 + (col & 15);
    float* r = rec + static_cast<size_t>(idx) * kColF;
    uint64_t cand[2];
    column_candidates(nb_index, bmask, cr, ci, x, z, cand);
    for (int k = 0; k < mctgen::kColNoises; ++k)
        r[k] = ((cand[0] & mctgen::kColNoiseBiomes[k][0]) | (cand[1] & mctgen::kColNoiseBiomes[k][1])) != 0
                   ? mctgen::surf_column_noise(seeds[cr.seed].oct, static_cast<double>(x), static_cast<double>(z), k)
                   : 0.0f;
    r[kColSnow] = snow_noise(M, x, z);
    r[kColPatch] = frozen_patch(M, x, z) ? 1.0f : 0.0f;
}

// ---- zoom down one block column -------------------------------------------------------------
// BiomeManager.getBiome for the blocks of one column (fixed x, z): the fiddle indices of the
// 8 corner quarts depend on the quart row only, so the column keeps those of the two rows the
// block lies between (4 corners each, packed mx | my << 10 | mz << 20) and recomputes the LCG
// chains of 4 corners when the walk moves down one quart row. The fiddled distances are the
// plain FP64 ones of zoom_pick, in the same order.
struct ZoomColumn {
    int64_t seed;
    int px, pz, fx, fz, py = INT32_MIN;
    uint32_t zc[8];

    MCT_HD MCT_FI void init(int64_t zoom_seed, int x, int z) {
        seed = zoom_seed;
        px = (x - 2) >> 2;
        pz = (z - 2) >> 2;
        fx = (x - 2) & 3;
        fz = (z - 2) & 3;
        py = INT32_MIN;
    }
    MCT_HD MCT_FI void layer(int cy, uint32_t* L) const {
        for (int k = 0; k < 4; ++k) {
            int mx, my, mz;
            corner_fiddles(seed, px + (k >> 1), cy, pz + (k & 1), mx, my, mz);
            L[k] = static_cast<uint32_t>(mx) | (static_cast<uint32_t>(my) << 10) | (static_cast<uint32_t>(mz) << 20);
        }
    }
    // the quart picked for block y of the column (quart x, y, z)
    MCT_HD MCT_FI void pick(int y, int& qx, int& qy, int& qz) {
        const int ay = y - 2, ry = ay >> 2, fy = ay & 3;
        if (ry != py) {
            if (ry == py - 1) {
                for (int k = 0; k < 4; ++k) zc[4 + k] = zc[k];
                layer(ry, zc);
            } else {
                layer(ry, zc);
                layer(ry + 1, zc + 4);
            }
            py = ry;
        }
        int best = 0;
        double best_d = u2d(0x7FF0000000000000ull);
        for (int i = 0; i < 8; ++i) {
            const bool xf = (i & 4) != 0, yf = (i & 2) != 0, zf = (i & 1) != 0;
            const uint32_t pk = zc[(yf ? 4 : 0) + (xf ? 2 : 0) + (zf ? 1 : 0)];
            const int mx = static_cast<int>(pk & 1023u), my = static_cast<int>((pk >> 10) & 1023u), mz = static_cast<int>(pk >> 20);
            const double d = (zoom_sq(fz, zf, mz) + zoom_sq(fy, yf, my)) + zoom_sq(fx, xf, mx);
            if (best_d > d) {
                best = i;
                best_d = d;
            }
        }
        qx = (best & 4) ? px + 1 : px;
        qy = (best & 2) ? ry + 1 : ry;
        qz = (best & 1) ? pz + 1 : pz;
    }
    // Candidates of the last pick within `margin` of its distance (1 unless a near tie): the
    // picks an FP32 distance filter with that margin could not decide.
    MCT_HD MCT_FI int candidates(int y, double margin) const {
        const int fy = (y - 2) & 3;
        double d[8], dmin = u2d(0x7FF0000000000000ull);
        for (int i = 0; i < 8; ++i) {
            const bool xf = (i & 4) != 0, yf = (i & 2) != 0, zf = (i & 1) != 0;
            const uint32_t pk = zc[(yf ? 4 : 0) + (xf ? 2 : 0) + (zf ? 1 : 0)];
            d[i] = (zoom_sq(fz, zf, static_cast<int>(pk >> 20)) + zoom_sq(fy, yf, static_cast<int>((pk >> 10) & 1023u))) + zoom_sq(fx, xf, static_cast<int>(pk & 1023u));
            dmin = d[i] < dmin ? d[i] : dmin;
        }
        int n = 0;
        for (int i = 0; i < 8; ++i) n += d[i] <= dmin + margin ? 1 : 0;
        return n;
    }
};

// ---- material system: one column ----------------------------------------------------------
struct SurfArgs {
    const SeedDev* seeds;
    const ModelDev* M;
    const ChunkRef* chunks;
    uint8_t* blocks;
    const int* nb_index;        // n x 9 biome-chunk indices
    const uint8_t* bbiomes;     // biome chunks x kQPer
    const float* corners;
    const int* colmap;
    const float* prelim;        // n x 4 chunk-corner preliminary surface values
    const float* colrec;        // n x 256 column records (kColF floats, column = lx * 16 + lz)
    uint8_t* ore;               // n x 256 x kOreRowsPad ore-vein outcomes (ore_precompute), or null
    const uint64_t* bmask;      // biome chunks x 16 quart columns x 2: biomes of each quart column, or null
    // n x 24 x 256 block-class words: word w of column t = z * 16 + x holds the classes
    // (block_class, 2 bits) of blocks w * 16 .. w * 16 + 15 above minY, the lowest at bit 0.
    // Written by the fill, kept current by the badlands extension; read by the LDS walk.
    uint32_t* clsw;
};

// ---- ore veins ahead of the surface rules -------------------------------------------------
// An ore-vein rule k is only reached for blocks with y in [kOreYLo[k], kOreYHi[k]) (the plain
// y thresholds on its paths through the rules). Its outcome at a block depends on the block
// alone (position, the interpolated ore channels, the seed's ore random), so it is computed
// for every block of that range in a dense pass before the surface walk, one byte per block:
// 0xFF (none) or the state.
// Per block column (chunk * 256 + lz * 16 + lx) kOreRowsPad bytes: the rows of rule 0, then
// rule 1, ..., so that the walk down a column reads them 16 at a time.
constexpr int kOreRules = mctgen::kNumOreRules;
constexpr int ore_rows(int k) { return mctgen::kOreYHi[k] > mctgen::kOreYLo[k] ? mctgen::kOreYHi[k] - mctgen::kOreYLo[k] : 0; }
constexpr int ore_cells(int k) {
    return ore_rows(k) > 0 ? ((mctgen::kOreYHi[k] - 1 - kMinY) >> 3) - ((mctgen::kOreYLo[k] - kMinY) >> 3) + 1 : 0;
}
constexpr int ore_row_offset(int k) {
    int s = 0;
    for (int j = 0; j < k; ++j) s += ore_rows(j);
    return s;
}
constexpr int kOreRows = ore_row_offset(kOreRules);
constexpr int kOreRowsPad = (kOreRows + 15) & ~15;

// Block classes the surface walk distinguishes: solid, air, water (the water state itself) and
// other fluids.
constexpr int kClsSolid = 0, kClsAir = 1, kClsWater = 2, kClsFluid = 3;
MCT_HD constexpr int block_class(int s) {
    if ((mctgen::kAirMask >> s) & 1u) return kClsAir;
    if (s == mctgen::kWater) return kClsWater;
    return ((mctgen::kFluidMask >> s) & 1u) ? kClsFluid : kClsSolid;
}
// block_class of palette indices 0..31, 2 bits each, so that a state known to be below 32 (the
// fill's outcomes: the default block, air and the fluids) takes one shift and one mask.
constexpr uint64_t kClassTable = [] {
    uint64_t t = 0;
    for (int s = 0; s < 32; ++s) t |= static_cast<uint64_t>(block_class(s)) << (2 * s);
    return t;
}();
MCT_HD MCT_FI int block_class_small(int s) { return static_cast<int>((kClassTable >> (2 * s)) & 3u); }
static_assert(mctgen::kDefaultBlock < 32 && mctgen::kAir < 32 && mctgen::kWater < 32 && mctgen::kLava < 32 && mctgen::kDefaultFluid < 32,
              "the fill's outcomes index the 32-entry class table");


struct Surf {
    const SurfArgs* A;
    const Octave* O;
    const SeedDev* S;
    int ci, cx, cz, lx, lz;
    uint8_t* blocks;
    uint8_t* col;
    int stride;
    uint32_t MCT_LDS* cls;   // the column's block classes in LDS (word w at cls[w * 64]), or null
    int* hm;
    const int* cmap;
    const float* rec;
    ZoomColumn zoom;
    int bc_py, bc_biome;   // quart row whose 8 zoom candidates were checked; their common biome or -1
    int bc_first;          // the first candidate's biome (a valid placeholder)
    int qoff[4];           // zoom_columns
    uint32_t row_lo, row_hi;   // candidate biomes of quart rows bc_py and bc_py + 1
    uint64_t bc_any[2];    // biomes among that row's 8 candidates (bit b of word b >> 6)
    uint64_t col_any[2];   // biomes of the 4 candidate quart columns at any height (all bits: unknown)
    int ore_cy;            // cell row of the cached ore-channel interpolation
    float ore_lo[kCh], ore_step[kCh];
    float prelim;
    int x, y, z;
    int grad_x, grad_z, surface_depth;
    int stone_above, stone_below, water_height;
    bool sec_set, biome_set, min_set, ore_set;
    bool need_biome;       // a rule needed the exact biome before it was known (see surface_walk)
    double secondary;
    int biome_v, min_surface;
    float ore_ch[kCh];

    MCT_HD MCT_FI int quart_biome(int qx, int qy, int qz) const {
        const int dx = (qx >> 2) - cx, dz = (qz >> 2) - cz;
        const int qmin = kMinY >> 2;
        const int cqy = clampi(qy, qmin, qmin + kQH - 1) - qmin;
        const size_t bc = static_cast<size_t>(A->nb_index[ci * 9 + (dz + 1) * 3 + dx + 1]);
        return A->bbiomes[bc * kQPer + (cqy * 4 + (qz & 3)) * 4 + (qx & 3)];
    }
    // BiomeManager.getBiome at (x, by, z) of this column. The zoom picks one of the 8 quarts
    // around (x-2, y-2, z-2) >> 2, which depend on the quart row only: where all 8 hold the
    // same biome, that is the answer for the 4 blocks of the row, whichever quart the fiddled
    // distances would pick.
    // The 4 quart columns the zoom candidates lie in ((px, pz) + (dx, dz), j = dx * 2 + dz) as
    // offsets into the biome chunks' quarts, set once per column; a quart row's 4 biomes packed
    // in a word, byte j.
    MCT_HD MCT_FI void zoom_columns() {
        col_any[0] = col_any[1] = A->bmask ? 0 : ~0ull;
        for (int j = 0; j < 4; ++j) {
            const int qx = zoom.px + (j >> 1), qz = zoom.pz + (j & 1);
            const int dx = (qx >> 2) - cx, dz = (qz >> 2) - cz;
            const int bc = A->nb_index[ci * 9 + (dz + 1) * 3 + dx + 1];
            qoff[j] = bc * kQPer + (qz & 3) * 4 + (qx & 3);
            if (A->bmask) {
                const uint64_t* m = A->bmask + (static_cast<size_t>(bc) * 16 + (qz & 3) * 4 + (qx & 3)) * 2;
                col_any[0] |= m[0];
                col_any[1] |= m[1];
            }
        }
    }
    MCT_HD MCT_FI uint32_t zoom_quart_row(int qy) const {
        const int qmin = kMinY >> 2;
        const int o = (clampi(qy, qmin, qmin + kQH - 1) - qmin) * 16;
        uint32_t w = 0;
        for (int j = 0; j < 4; ++j) w |= static_cast<uint32_t>(A->bbiomes[qoff[j] + o]) << (8 * j);
        return w;
    }
    // The 8 zoom candidates of block row `by` (quart rows py and py + 1 of the 4 columns): walking
    // down, each new row reuses the previous lower row.
    MCT_HD MCT_FI void zoom_row(int by) {
        const int py = (by - 2) >> 2;
        if (py == bc_py) return;
        if (py == bc_py - 1) {
            row_hi = row_lo;
            row_lo = zoom_quart_row(py);
        } else {
            row_lo = zoom_quart_row(py);
            row_hi = zoom_quart_row(py + 1);
        }
        bc_py = py;
        const int b0 = static_cast<int>(row_lo & 0xFFu);
        uint64_t any0 = 0, any1 = 0;
        bool same = true;
        for (int i = 0; i < 8; ++i) {
            const uint32_t row = ((i >> 1) & 1) ? row_hi : row_lo;
            const int b = static_cast<int>((row >> (8 * (((i >> 2) & 1) * 2 + (i & 1)))) & 0xFFu);
            same = same && b == b0;
            if (b < 64) any0 |= 1ull << b;
            else any1 |= 1ull << (b - 64);
        }
        bc_biome = same ? b0 : -1;
        bc_first = b0;
        bc_any[0] = any0;
        bc_any[1] = any1;
    }
    MCT_HD MCT_FI int biome_y(int by) {
        zoom_row(by);
        if (bc_biome >= 0) return bc_biome;
        int qx, qy, qz;
        zoom.pick(by, qx, qy, qz);
        return quart_biome(qx, qy, qz);
    }
    // The block's biome inside the rules. The zoom is not run here (its code would be inlined
    // at every biome test of the rule tree): where it is needed and not yet known, the rule
    // evaluation flags it, carries on with a placeholder, and the walk runs the zoom once and
    // evaluates the rules again (they have no effects beyond per-block memos).
    MCT_HD MCT_FI int biome() {
        if (!biome_set) {
            zoom_row(y);
            need_biome = true;
            return bc_first;
        }
        return biome_v;
    }
    // Whether the block's biome is in the set {b : bit b of (lo, hi)}. The zoom picks one of the
    // row's 8 candidate quarts (zoom_row, loaded here when first needed), so where the set holds
    // all or none of them the answer needs no zoom; where it holds all or none of the biomes of
    // the 4 candidate columns at any height, not even the row.
    MCT_HD MCT_FI bool biome_in(uint64_t lo, uint64_t hi) {
        if (!biome_set) {
            if (((col_any[0] & lo) | (col_any[1] & hi)) == 0) return false;
            if (((col_any[0] & ~lo) | (col_any[1] & ~hi)) == 0) return true;
            zoom_row(y);
            if (((bc_any[0] & lo) | (bc_any[1] & hi)) == 0) return false;
            if (((bc_any[0] & ~lo) | (bc_any[1] & ~hi)) == 0) return true;
        }
        const int b = biome();
        return ((b < 64 ? (lo >> b) : (hi >> (b - 64))) & 1ull) != 0;
    }
    MCT_HD MCT_FI float col_noise(int k) const { return rec[k]; }
    MCT_HD MCT_FI float temperature(int b, int by) const { return biome_temperature_col(*A->M, b, by, rec[kColPatch] != 0.0f, rec[kColSnow]); }
    MCT_HD MCT_FI bool vertical_gradient(int k, int true_at, int false_at) const {
        if (y <= true_at) return true;
        if (y >= false_at) return false;
        const double t = (static_cast<double>(y) - true_at) / static_cast<double>(false_at - true_at);
        const double probability = 1.0 + t * (0.0 - 1.0);
        Xoro r = at(S->grad[k], x, y, z);
        return static_cast<double>(next_float(r)) < probability;
    }
    MCT_HD MCT_FI double surface_secondary() {
        if (!sec_set) { secondary = static_cast<double>(rec[mctgen::kColSecondary]); sec_set = true; }
        return secondary;
    }
    MCT_HD MCT_FI bool stone_depth(int offset, bool add, int range, bool ceiling) {
        const int stone = ceiling ? stone_below : stone_above;
        const int sd = add ? surface_depth : 0;
        int sec = 0;
        if (range != 0) {
            const double t = (surface_secondary() - -1.0) / (1.0 - -1.0);
            sec = d2i(0.0 + t * (static_cast<double>(range) - 0.0));
        }
        return stone <= 1 + offset + sd + sec;
    }
    MCT_HD MCT_FI bool cold_enough_to_snow() { return !(temperature(biome(), y) >= 0.15f); }
    MCT_HD MCT_FI int min_surface_level() {
        if (!min_set) { min_surface = mfloor(static_cast<double>(prelim)) + surface_depth - 8; min_set = true; }
        return min_surface;
    }
    // The remaining SurfaceRules conditions (cpu/src/terrain.cpp:454-746, Surface::test).
    // NoiseThreshold: the noise value, widened to double, in [lo, hi].
    MCT_HD MCT_FI bool noise_in(float v, double lo, double hi) const {
        const double d = static_cast<double>(v);
        return d >= lo && d <= hi;
    }
    // YConditionSource: y (+ the stone depth above) at or above anchor + surface depth * m.
    MCT_HD MCT_FI bool y_above(int anchor, int m, bool add_stone) const { return y + (add_stone ? stone_above : 0) >= anchor + surface_depth * m; }
    // WaterConditionSource: no water above, or y (+ stone depth) at or above the water height
    // + offset + surface depth * m.
    MCT_HD MCT_FI bool water_above(int offset, int m, bool add_stone) const {
        return water_height == INT32_MIN || y + (add_stone ? stone_above : 0) >= water_height + offset + surface_depth * m;
    }
    MCT_HD MCT_FI bool steep() const { return grad_x <= -4 || grad_z >= 4; }
    MCT_HD MCT_FI bool hole() const { return surface_depth <= 0; }
    MCT_HD MCT_FI bool above_preliminary_surface() { return y >= min_surface_level(); }
    MCT_HD MCT_FI int band() const {
        const int offset = jroundf(rec[mctgen::kColClay] * 4.0f);
        return S->clay[(y + offset + 192) % 192];
    }
    // Ore veins (OreVeinifier as the model's roots express it).
    MCT_HD MCT_FI int ore_vein(int k) {
        if (!ore_set) {
            // the ore evaluators read channels from kOreChanFirst on; the cell's z / x lerps
            // are kept while the walk stays in one cell row
            const int yy = y - kMinY, cy = yy >> 3;
            if (cy != ore_cy) {
                ore_cy = cy;
                const int cx = lx >> 2, xi = lx & 3, cz = lz >> 2, zi = lz & 3;
                const size_t u00 = cmap[cz * 5 + cx], u10 = cmap[cz * 5 + cx + 1], u01 = cmap[(cz + 1) * 5 + cx], u11 = cmap[(cz + 1) * 5 + cx + 1];
                for (int k = mctgen::kOreChanFirst; k < kCh; ++k) {
                    const float* a = A->corners + (u00 * kCh + k) * kRows + cy;
                    const float* b = A->corners + (u10 * kCh + k) * kRows + cy;
                    const float* cc = A->corners + (u01 * kCh + k) * kRows + cy;
                    const float* d = A->corners + (u11 * kCh + k) * kRows + cy;
                    cell_lo_step(a[0], b[0], a[1], b[1], cc[0], d[0], cc[1], d[1], xi, zi, 0.25f, 0.125f, ore_lo[k], ore_step[k]);
                }
            }
            for (int k = mctgen::kOreChanFirst; k < kCh; ++k) ore_ch[k] = cell_at(ore_lo[k], ore_step[k], yy & 7);
            ore_set = true;
        }
        const float density = mctgen::eval_ore(O, x, y, z, ore_ch, 2 * k);
        if (density <= 0.0f) return -1;
        Xoro r = at(S->ore, x, y, z);
        if (next_float(r) > density) return -1;
        const float richness = mctgen::eval_ore(O, x, y, z, ore_ch, 2 * k + 1);
        if (next_float(r) < richness && mctgen::eval_point(O, x, y, z, 5 + k) < 0.0f)
            return next_float(r) < A->M->ore_raw[k] ? A->M->ore_state[k][1] : A->M->ore_state[k][0];
        return A->M->ore_state[k][2];
    }

    // the column's blocks: col[(y - minY) * stride] (the chunk in global memory, stride 256)
    MCT_HD MCT_FI size_t bidx(int by) const { return static_cast<size_t>(by - kMinY) * static_cast<size_t>(stride); }
    MCT_HD MCT_FI int get(int by) const { return (by < kMinY || by >= kMinY + kHeight) ? mctgen::kAir : col[bidx(by)]; }
    MCT_HD MCT_FI bool air(int s) const { return ((mctgen::kAirMask >> s) & 1u) != 0; }
    MCT_HD MCT_FI bool fluid(int s) const { return ((mctgen::kFluidMask >> s) & 1u) != 0; }
    // setBlockState + Heightmap.update(WORLD_SURFACE_WG) on the column in memory; `write` false
    // moves only the height
    MCT_HD MCT_FI void set(int by, int state, bool write) {
        if (by < kMinY || by >= kMinY + kHeight) return;
        if (write) col[bidx(by)] = static_cast<uint8_t>(state);
        const int first = *hm;
        if (by <= first - 2) return;
        if (!air(state)) {
            if (by >= first) *hm = by + 1;
        } else if (first - 1 == by) {
            for (int yy = by - 1; yy >= kMinY; --yy)
                if (!air(get(yy))) { *hm = yy + 1; return; }
            *hm = kMinY;
        }
    }
};

// The surface state of the LDS walk: the column record in LDS, ore veins read from the dense
// precompute (ore_precompute).
struct SurfGap : Surf {
    const float MCT_LDS* lrec;   // the column record staged in LDS (value k at lrec[k * 64])
    // the ore outcomes of 16 rows of the column (rows ore_wb .. ore_wb + 15), kept while the
    // walk moves down through them
    int ore_wb;
    uint32_t ore_w[4];
    MCT_HD MCT_FI int ore_vein(int k) {
        const int r = ore_row_offset(k) + (y - mctgen::kOreYLo[k]);
        if ((r & ~15) != ore_wb) {
            ore_wb = r & ~15;
            struct alignas(16) Row16 { uint32_t w[4]; };
            const Row16 w = *reinterpret_cast<const Row16*>(A->ore + (static_cast<size_t>(ci) * 256 + lz * 16 + lx) * kOreRowsPad + ore_wb);
            for (int j = 0; j < 4; ++j) ore_w[j] = w.w[j];
        }
        const int q = (r >> 2) & 3;
        const uint32_t w = q == 0 ? ore_w[0] : q == 1 ? ore_w[1] : q == 2 ? ore_w[2] : ore_w[3];
        const int v = static_cast<int>((w >> ((r & 3) * 8)) & 0xFFu);
        return v == 0xFF ? -1 : v;
    }
    MCT_HD MCT_FI float col_noise(int k) const { return lrec[k * 64]; }
    MCT_HD MCT_FI int band() const {
        const int offset = jroundf(lrec[mctgen::kColClay * 64] * 4.0f);
        return S->clay[(y + offset + 192) % 192];
    }
};
static_assert(mctgen::kPaletteSize < 0xFF, "0xFF marks no ore vein in the precomputed bytes");

// Ore-vein outcomes of one block column (idx = chunk * 256 + lz * 16 + lx), rows in order, as
// Surf::ore_vein gives them: per cell row the z / x lerps of the ore channels, then y as the
// running sum of the step from the cell's bottom. The bytes are stored 16 at a time. Blocks
// that pass the richness test need the filler-gap noise: they are marked in a row mask (with
// the raw-ore draw, taken at once from the block's own random) and evaluated afterwards, every
// lane of the wave taking its next marked row together, until none is left.
constexpr int kOreMaskWords = (kOreRowsPad + 31) / 32;
MCT_HD inline void ore_precompute(const SurfArgs& A, int idx) {
    const int col = idx & 255, lx = col & 15, lz = col >> 4, c = idx >> 8;
    const ChunkRef cr = A.chunks[c];
    const SeedDev& S = A.seeds[cr.seed];
    const ModelDev& M = *A.M;
    const int* cmap = A.colmap + static_cast<size_t>(c) * kCols;
    const int cx = lx >> 2, xi = lx & 3, cz = lz >> 2, zi = lz & 3;
    const size_t u00 = cmap[cz * 5 + cx], u10 = cmap[cz * 5 + cx + 1], u01 = cmap[(cz + 1) * 5 + cx], u11 = cmap[(cz + 1) * 5 + cx + 1];
    const int x = cr.x * 16 + lx, z = cr.z * 16 + lz;
    uint8_t* out = A.ore + static_cast<size_t>(idx) * kOreRowsPad;
    struct alignas(16) Row16 { uint32_t w[4]; };
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0, acc = 0;   // the last 4 complete words, the open one
    uint32_t gap[kOreMaskWords], raw[kOreMaskWords];   // rows waiting for the gap noise; their raw draw
    for (int j = 0; j < kOreMaskWords; ++j) gap[j] = raw[j] = 0;
    int r = 0;
    for (int k = 0; k < kOreRules; ++k) {
        if (ore_rows(k) == 0) continue;
        const int cy0 = (mctgen::kOreYLo[k] - kMinY) >> 3;
        for (int cy = cy0; cy < cy0 + ore_cells(k); ++cy) {
            float ch[kCh], step[kCh];
            for (int q = 0; q < kCh; ++q) ch[q] = step[q] = 0.0f;
            for (int q = mctgen::kOreChanFirst; q < kCh; ++q) {
                const float* a = A.corners + (u00 * kCh + q) * kRows + cy;
                const float* b = A.corners + (u10 * kCh + q) * kRows + cy;
                const float* cc = A.corners + (u01 * kCh + q) * kRows + cy;
                const float* d = A.corners + (u11 * kCh + q) * kRows + cy;
                float lo;
                cell_lo_step(a[0], b[0], a[1], b[1], cc[0], d[0], cc[1], d[1], xi, zi, 0.25f, 0.125f, lo, step[q]);
                ch[q] = lo + step[q] * 0.0f;
            }
            for (int yi = 0; yi < 8; ++yi) {
                const int y = kMinY + cy * 8 + yi;
                if (y >= mctgen::kOreYLo[k] && y < mctgen::kOreYHi[k]) {
                    uint32_t v = 0xFF;
                    const float density = mctgen::eval_ore(S.oct, x, y, z, ch, 2 * k);
                    if (density > 0.0f) {
                        Xoro rnd = at(S.ore, x, y, z);
                        if (!(next_float(rnd) > density)) {
                            const float richness = mctgen::eval_ore(S.oct, x, y, z, ch, 2 * k + 1);
                            v = static_cast<uint32_t>(M.ore_state[k][2]);
                            if (next_float(rnd) < richness) {
                                const bool rw = next_float(rnd) < M.ore_raw[k];
                                for (int j = 0; j < kOreMaskWords; ++j) {
                                    if ((r >> 5) == j) {
                                        gap[j] |= 1u << (r & 31);
                                        if (rw) raw[j] |= 1u << (r & 31);
                                    }
                                }
                            }
                        }
                    }
                    acc |= v << ((r & 3) * 8);
                    if ((r & 3) == 3) { w0 = w1; w1 = w2; w2 = w3; w3 = acc; acc = 0; }
                    if ((r & 15) == 15) *reinterpret_cast<Row16*>(out + (r - 15)) = Row16{{w0, w1, w2, w3}};
                    ++r;
                }
                for (int q = mctgen::kOreChanFirst; q < kCh; ++q) ch[q] += step[q];
            }
        }
    }
    if (r & 15) {   // the last, partial 16 rows
        if (r & 3) { w0 = w1; w1 = w2; w2 = w3; w3 = acc; }
        for (int j = ((r & 15) + 3) >> 2; j < 4; ++j) { w0 = w1; w1 = w2; w2 = w3; w3 = 0; }
        *reinterpret_cast<Row16*>(out + (r & ~15)) = Row16{{w0, w1, w2, w3}};
    }
    for (;;) {
        int row = -1;
        for (int j = kOreMaskWords - 1; j >= 0 && row < 0; --j)
            if (gap[j]) row = j * 32 + 31 - __builtin_clz(gap[j]);
#if defined(__HIP_DEVICE_COMPILE__)
        if (__builtin_amdgcn_ballot_w64(row >= 0) == 0) break;
#else
        if (row < 0) break;
#endif
        if (row >= 0) {
            int k = 0;
            while (k < kOreRules - 1 && row >= ore_row_offset(k + 1)) ++k;
            const int y = mctgen::kOreYLo[k] + row - ore_row_offset(k);
            bool rw = false;
            for (int j = 0; j < kOreMaskWords; ++j) {
                if ((row >> 5) == j) {
                    gap[j] &= ~(1u << (row & 31));
                    rw = ((raw[j] >> (row & 31)) & 1u) != 0;
                }
            }
            if (mctgen::eval_point(S.oct, x, y, z, 5 + k) < 0.0f) out[row] = static_cast<uint8_t>(rw ? M.ore_state[k][1] : M.ore_state[k][0]);
        }
    }
}

MCT_HD MCT_FI void surf_init(Surf& c, const SurfArgs& A, int ci, int lx, int lz, int* hm) {
    c.A = &A;
    const ChunkRef cr = A.chunks[ci];
    c.S = &A.seeds[cr.seed];
    c.O = c.S->oct;
    c.cx = cr.x;
    c.cz = cr.z;
    c.lx = lx;
    c.lz = lz;
    c.blocks = A.blocks + static_cast<size_t>(ci) * kPer;
    c.col = c.blocks + lz * 16 + lx;
    c.stride = 256;
    c.cls = nullptr;
    c.hm = hm;
    c.ci = ci;
    c.rec = A.colrec + (static_cast<size_t>(ci) * 256 + static_cast<size_t>(lx * 16 + lz)) * kColF;
    c.cmap = A.colmap + static_cast<size_t>(ci) * kCols;
    const float* pc = A.prelim + static_cast<size_t>(ci) * 4;
    // chunk_surface_level: InterpolatedFunction 16 x 1 over the chunk's 4 corners
    c.prelim = cell_fill(pc[0], pc[1], pc[0], pc[1], pc[2], pc[3], pc[2], pc[3], lx, 0, lz, 1.0f / 16.0f, 1.0f / 1.0f);
    c.sec_set = c.biome_set = c.min_set = c.ore_set = c.need_biome = false;
    c.x = cr.x * 16 + lx;
    c.z = cr.z * 16 + lz;
    c.y = 0;
    c.zoom.init(c.S->zoom_seed, c.x, c.z);
    c.zoom_columns();
    c.bc_py = INT32_MIN;
    c.bc_biome = -1;
    c.ore_cy = -1;
}

MCT_HD MCT_FI void surf_update_xz(Surf& c, int gx, int gz) {
    c.grad_x = gx;
    c.grad_z = gz;
    const double n = static_cast<double>(c.rec[mctgen::kColSurface]);
    Xoro r = at(c.S->random, c.x, 0, c.z);
    c.surface_depth = d2i(n * 2.75 + 3.0 + next_double(r) * 0.25);
    c.sec_set = false;
    c.min_set = false;
}

// MaterialSystem eroded-badlands extension (writes blocks and moves the height).
MCT_HD MCT_FI bool eroded_badlands(Surf& c, int height) {
    const double bx = c.x, bz = c.z;
    const double a = fabs(mctgen::surf_badlands_surface(c.O, bx, 0.0, bz) * 8.25);
    const double b = static_cast<double>(static_cast<float>(mctgen::surf_badlands_pillar(c.O, bx * 0.2, 0.0, bz * 0.2)) * 15.0f);
    const double pillar = jmind(a, b);
    if (pillar <= 0.0) return false;
    const double floor_v = fabs(mctgen::surf_badlands_pillar_roof(c.O, bx * 0.75, 0.0, bz * 0.75) * 1.5);
    const double top = 64.0 + jmind(pillar * pillar * 2.5, ceil(floor_v * 50.0) + 24.0);
    const int start_y = mfloor(top);
    if (height > start_y) return false;
    for (int yy = start_y; yy >= kMinY; --yy) {
        const int old = c.get(yy);
        if (old == mctgen::kDefaultBlock) break;
        if ((mctgen::kWaterMask >> old) & 1u) return false;
    }
    bool wrote = false;
    for (int yy = start_y; yy >= kMinY && c.air(c.get(yy)); --yy) {
        c.set(yy, mctgen::kDefaultBlock, true);
        wrote = true;
    }
    return wrote;
}

// What the walk and the frozen-ocean extension test of a block: its class (kClsSolid / kClsAir
// / kClsWater / kClsFluid). L: from the column's LDS class words (2 bits per block, 16 blocks
// per word, Surf::cls), else from the block in memory.
template <bool L>
MCT_HD MCT_FI int block_at(const Surf& c, int by) {
    if (by < kMinY || by >= kMinY + kHeight) return kClsAir;
    if constexpr (L) {
        const int r = by - kMinY;
        return static_cast<int>((c.cls[(r >> 4) * 64] >> ((r & 15) * 2)) & 3u);
    } else {
        return block_class(c.col[c.bidx(by)]);
    }
}

// Surf::set, keeping the LDS class words in step when L.
template <bool L>
MCT_HD MCT_FI void block_set(Surf& c, int by, int state, bool write) {
    if constexpr (L) {
        if (by < kMinY || by >= kMinY + kHeight) return;
        if (write) {
            c.col[c.bidx(by)] = static_cast<uint8_t>(state);
            const int r = by - kMinY, sh = (r & 15) * 2;
            uint32_t MCT_LDS* w = c.cls + (r >> 4) * 64;
            *w = (*w & ~(3u << sh)) | (static_cast<uint32_t>(block_class(state)) << sh);
        }
        const int first = *c.hm;
        if (by <= first - 2) return;
        if (!c.air(state)) {
            if (by >= first) *c.hm = by + 1;
        } else if (first - 1 == by) {
            for (int yy = by - 1; yy >= kMinY; --yy)
                if (block_at<true>(c, yy) != kClsAir) { *c.hm = yy + 1; return; }
            *c.hm = kMinY;
        }
    } else {
        c.set(by, state, write);
    }
}

// MaterialSystem frozen-ocean extension; with write == false only the height moves.
template <bool L>
MCT_HD MCT_FI void frozen_ocean(Surf& c, int min_surface, int biome, int height, bool write) {
    const double bx = c.x, bz = c.z;
    const double a = fabs(mctgen::surf_iceberg_surface(c.O, bx, 0.0, bz) * 8.25);
    const double b = static_cast<double>(static_cast<float>(mctgen::surf_iceberg_pillar(c.O, bx * 1.28, 0.0, bz * 1.28)) * 15.0f);
    const double iceberg = jmind(a, b);
    if (iceberg <= 1.8) return;
    const double roof = fabs(mctgen::surf_iceberg_pillar_roof(c.O, bx * 1.17, 0.0, bz * 1.17) * 1.5);
    double top = jmind(iceberg * iceberg * 1.2, ceil(roof * 40.0) + 14.0);
    if (c.temperature(biome, kSea) > 0.1f) top -= 2.0;
    if (top <= 2.0) return;
    const double bottom = static_cast<double>(kSea) - top - 7.0;
    top += static_cast<double>(kSea);
    Xoro r = at(c.S->random, c.x, 0, c.z);
    const int max_snow = 2 + next_int(r, 4);
    const int min_snow_height = kSea + 18 + next_int(r, 10);
    int snow = 0;
    const int itop = d2i(top), ibottom = d2i(bottom);
    for (int yy = jmaxi(height, itop + 1); yy >= min_surface; --yy) {
        const int cur = block_at<L>(c, yy);
        bool place = false;
        if (cur == kClsAir && yy < itop && next_double(r) > 0.01) place = true;
        else if (cur == kClsWater && yy > ibottom && yy < kSea && next_double(r) > 0.15) place = true;
        if (place) {
            if (snow <= max_snow && yy > min_snow_height) {
                block_set<L>(c, yy, mctgen::kSnowBlock, write);
                ++snow;
            } else {
                block_set<L>(c, yy, mctgen::kPackedIce, write);
            }
        }
    }
}

// The material-rule walk of one column from `height` down, then the frozen-ocean extension.
template <bool L, class S>
MCT_HD MCT_FI void surface_walk(S& c, int height, int surface_biome, int starting) {
    int stone_above = 0, water_height = INT32_MIN, next_ceiling = INT32_MAX;
    uint32_t word = 0;   // the LDS class word of the current 16 blocks (L)
    int word_r = -1;
    for (int yy = height; yy >= kMinY; --yy) {
        int old;
        if constexpr (L) {
            const int r = yy - kMinY;
            if (r >= kHeight) {
                old = kClsAir;
            } else {
                if ((r >> 4) != word_r) {
                    word_r = r >> 4;
                    word = c.cls[word_r * 64];
                }
                old = static_cast<int>((word >> ((r & 15) * 2)) & 3u);
            }
        } else {
            old = block_at<L>(c, yy);
        }
        if (old == kClsAir) {
            stone_above = 0;
            water_height = INT32_MIN;
        } else if (old != kClsSolid) {
            if (water_height == INT32_MIN) water_height = yy + 1;
        } else {
            if (next_ceiling >= yy) {
                next_ceiling = kWayBelow;
                if constexpr (L) {
                    // the highest non-solid block below yy (below the world: air), 16 blocks per
                    // class word: bit 2i of (w | w >> 1) & 0x55555555 marks block i non-solid
                    next_ceiling = kMinY;
                    const int r = yy - 1 - kMinY;
                    for (int wi = r >> 4; r >= 0 && wi >= 0; --wi) {
                        const uint32_t w = c.cls[wi * 64];
                        uint32_t ns = (w | (w >> 1)) & 0x55555555u;
                        if (wi == (r >> 4)) ns &= (2u << (2 * (r & 15))) - 1u;
                        if (ns) {
                            next_ceiling = kMinY + wi * 16 + (31 - __builtin_clz(ns)) / 2 + 1;
                            break;
                        }
                    }
                } else {
                    for (int la = yy - 1; la >= kMinY - 1; --la) {
                        if (block_at<L>(c, la) != kClsSolid) { next_ceiling = la + 1; break; }
                    }
                }
            }
            ++stone_above;
            c.stone_above = stone_above;
            c.stone_below = yy - next_ceiling + 1;
            c.water_height = water_height;
            c.y = yy;
            c.biome_set = false;
            c.ore_set = false;
            c.need_biome = false;
            int state = mctgen::apply_rules(c);
            if (c.need_biome) {
                c.biome_v = c.biome_y(yy);
                c.biome_set = true;
                state = mctgen::apply_rules(c);
            }
            if (state >= 0) block_set<L>(c, yy, state, true);
        }
    }
    if (surface_biome == mctgen::kBiomeFrozenOcean || surface_biome == mctgen::kBiomeDeepFrozenOcean)
        frozen_ocean<L>(c, c.min_surface_level(), surface_biome, starting, true);
}

// Pass 1 of a column: eroded-badlands extension (real) -> h1, frozen-ocean dry run -> h2.
MCT_HD MCT_FI void surface_pass1(const SurfArgs& A, const int* h0, int* h1, int* h2, int ci, int col) {
    const int lx = col >> 4, lz = col & 15;
    const int hidx = ci * 256 + col;
    int hm = h0[hidx];
    Surf c;
    surf_init(c, A, ci, lx, lz, &hm);
    const int starting = hm;
    const int sb = c.biome_y(starting);
    if (sb == mctgen::kBiomeErodedBadlands && eroded_badlands(c, starting) && A.clsw) {
        // the column's class words again (the extension turned air into stone)
        for (int w = 0; w < kHeight / 16; ++w) {
            uint32_t bits = 0;
            for (int j = 0; j < 16; ++j) bits |= static_cast<uint32_t>(block_class(c.col[c.bidx(kMinY + w * 16 + j)])) << (2 * j);
            A.clsw[(static_cast<size_t>(ci) * (kHeight / 16) + w) * 256 + lz * 16 + lx] = bits;
        }
    }
    h1[hidx] = hm;
    if (sb == mctgen::kBiomeFrozenOcean || sb == mctgen::kBiomeDeepFrozenOcean) {
        surf_update_xz(c, 0, 0);
        frozen_ocean<false>(c, c.min_surface_level(), sb, starting, false);
    }
    h2[hidx] = hm;
}

// Pass 2 of a column, assuming earlier columns (x-major, z-minor order) end at h2 and later
// ones are still at h0; the final height goes to hf (checked against h2 afterwards).
// L: the column's block classes are staged in LDS at cls (Surf::cls), its record at lrec.
template <bool L>
MCT_HD MCT_FI void surface_pass2(const SurfArgs& A, const int* h0, const int* h1, const int* h2, int* hf, int ci, int col,
                                 uint32_t MCT_LDS* cls = nullptr, const float MCT_LDS* lrec = nullptr) {
    const int lx = col >> 4, lz = col & 15;
    const int hidx = ci * 256 + col;
    auto H = [&](int ox, int oz) {
        const int j = ci * 256 + ox * 16 + oz;
        int v;
        if (ox == lx && oz == lz) v = h1[j];
        else if (ox < lx || (ox == lx && oz < lz)) v = h2[j];
        else v = h0[j];
        return v - 1;
    };
    int hm = h1[hidx];
    typename std::conditional<L, SurfGap, Surf>::type c;
    surf_init(c, A, ci, lx, lz, &hm);
    c.cls = cls;
    if constexpr (L) {
        c.lrec = lrec;
        c.ore_wb = -16;
    }
    const int starting = h0[hidx];
    const int sb = c.biome_y(starting);
    const int height = H(lx, lz) + 1;
    const int gx = H(lx + 1 < 15 ? lx + 1 : 15, lz) - H(lx - 1 > 0 ? lx - 1 : 0, lz);
    const int gz = H(lx, lz + 1 < 15 ? lz + 1 : 15) - H(lx, lz - 1 > 0 ? lz - 1 : 0);
    surf_update_xz(c, gx, gz);
    surface_walk<L>(c, height, sb, starting);
    hf[hidx] = hm;
}

// Whether column col of a chunk that failed the speculative check must be walked again in the
// exact order. A walk reads the other columns' heights only through the slope of its steep
// test (grad_x <= -4 || grad_z >= 4, from the x+-1 and z+-1 neighbours, clamped to the chunk);
// the speculative pass took earlier columns (x-major) at h2, later ones at h0 and the column
// itself at h1. hm holds the final heights of the columns already processed and h0 of the rest,
// so the live slope takes those, and h1 for the column itself (its own badlands extension runs
// first in either order). The same steep answer means the same walk, its blocks and height hf.
MCT_HD MCT_FI bool surface_replay_needed(const int* hm, const int* h0c, const int* h1c, const int* h2c, int col) {
    const int lx = col >> 4, lz = col & 15;
    const int xp = lx < 15 ? col + 16 : col, xm = lx > 0 ? col - 16 : col;
    const int zp = lz < 15 ? col + 1 : col, zm = lz > 0 ? col - 1 : col;
    auto spec = [&](int j) { return j == col ? h1c[j] : (j < col ? h2c[j] : h0c[j]); };
    auto live = [&](int j) { return j == col ? h1c[j] : hm[j]; };
    const bool s_spec = spec(xp) - spec(xm) <= -4 || spec(zp) - spec(zm) >= 4;
    const bool s_live = live(xp) - live(xm) <= -4 || live(zp) - live(zm) >= 4;
    return s_spec != s_live;
}

// One column in the exact sequential order, with a live height array hm[256] of the chunk
// (used for chunks whose speculative pass 2 saw a height change it did not assume).
MCT_HD MCT_FI void surface_exact_column(const SurfArgs& A, int* hm, int ci, int col) {
    const int lx = col >> 4, lz = col & 15;
    Surf c;
    surf_init(c, A, ci, lx, lz, hm + col);
    auto H = [&](int ox, int oz) { return hm[ox * 16 + oz] - 1; };
    const int starting = H(lx, lz) + 1;
    const int sb = c.biome_y(starting);
    if (sb == mctgen::kBiomeErodedBadlands) eroded_badlands(c, starting);
    const int height = H(lx, lz) + 1;
    const int gx = H(lx + 1 < 15 ? lx + 1 : 15, lz) - H(lx - 1 > 0 ? lx - 1 : 0, lz);
    const int gz = H(lx, lz + 1 < 15 ? lz + 1 : 15) - H(lx, lz - 1 > 0 ? lz - 1 : 0);
    surf_update_xz(c, gx, gz);
    surface_walk<false>(c, height, sb, starting);
}

}  // namespace mct
