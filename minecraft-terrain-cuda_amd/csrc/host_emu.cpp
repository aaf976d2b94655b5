// Host emulation of the GPU TERRAIN pipeline: the kernel bodies of pipeline.h and the
// generated evaluators, compiled for the CPU and run stage by stage in the same order and
// with the same data layouts as the device launches (the dedupe uses std::sort instead of
// the device radix sort). It shares the product's code paths, so it is a debugging and
// bisection tool for them (per stage, per value), not an independent check: the
// independent checks are the oracle and the reference port.
#include <algorithm>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "pipeline.h"

using namespace mct;

namespace {

thread_local std::string g_err;

// Operations of each stage of the last mcte_generate on this thread, named as the device's
// kernel groups (the build with MCT_COUNT_OPS counts; every other build records zeros).
struct StageOps {
    const char* name;
    uint64_t ops[4];   // FP64, FP32, INT32, memory reads
};
thread_local std::vector<StageOps> g_stages;

void stage_begin() {
    g_stages.clear();
#if defined(MCT_COUNT_OPS)
    g_tally = OpTally{};
#endif
}

// Everything counted since the previous mark belongs to stage `name`.
void stage_mark(const char* name) {
    StageOps st{name, {0, 0, 0, 0}};
#if defined(MCT_COUNT_OPS)
    st.ops[0] = g_tally.f64;
    st.ops[1] = g_tally.f32;
    st.ops[2] = g_tally.i32;
    st.ops[3] = g_tally.mem;
    g_tally = OpTally{};
#endif
    g_stages.push_back(st);
}

struct Emu {
    ModelDev M{};
    std::vector<TreeNodeDev> tree;
    std::vector<TreeNodeF> treef;
    std::vector<uint8_t> leaf_biome;
    std::vector<int32_t> leaf_node;
    std::vector<int32_t> leaf_rank;
    std::unique_ptr<SeedDev> seed;
    int64_t seed_value = 0;
    bool replay_all = false;
    bool has_seed = false;
};

void fill_model(Emu& e) {
    e.tree.resize(mctgen::kTreeNodes);
    for (int i = 0; i < mctgen::kTreeNodes; ++i) {
        const mctgen::TreeNode& t = mctgen::kTree[i];
        TreeNodeDev& d = e.tree[static_cast<size_t>(i)];
        for (int k = 0; k < 7; ++k) { d.lo[k] = t.lo[k]; d.hi[k] = t.hi[k]; }
        d.first = t.first;
        d.count = t.count;
        d.leaf = t.leaf;
        d.pad = 0;
    }
    e.treef = float_tree(e.tree);
    e.leaf_biome.assign(mctgen::kLeafBiome, mctgen::kLeafBiome + mctgen::kTreeLeaves);
    e.leaf_node.assign(mctgen::kLeafNode, mctgen::kLeafNode + mctgen::kTreeLeaves);
    ModelDev& M = e.M;
    M.tree = e.tree.data();
    M.treef = e.treef.empty() ? nullptr : e.treef.data();
    M.leaf_biome = e.leaf_biome.data();
    M.leaf_node = e.leaf_node.data();
    e.leaf_rank = leaf_dfs_rank(e.tree, mctgen::kTreeLeaves);
    M.leaf_rank = e.leaf_rank.data();
    if (!e.treef.empty()) column_groups(e.treef, M);
    std::memcpy(M.block_flags, mctgen::kBlockFlags, 64);
    std::memcpy(M.biome_temp, mctgen::kBiomeTemp, sizeof M.biome_temp);
    std::memcpy(M.biome_frozen, mctgen::kBiomeFrozen, sizeof M.biome_frozen);
    std::memcpy(M.simplex, mctgen::kSimplexPerm, sizeof M.simplex);
    for (int k = 0; k < 7; ++k) M.terracotta[k] = mctgen::kTerracotta[k];
    for (int k = 0; k < mctgen::kNumOreRules; ++k) {
        for (int j = 0; j < 3; ++j) M.ore_state[k][j] = mctgen::kOreState[k][j];
        M.ore_raw[k] = mctgen::kOreRawChance[k];
    }
}

void build_seed(Emu& e, int64_t seed) {
    if (e.has_seed && e.seed_value == seed) return;
    if (!e.seed) e.seed = std::make_unique<SeedDev>();
    SeedDev& S = *e.seed;
    uint8_t scratch[256];
    int blended_first[8];
    for (int b = 0; b < 8; ++b) blended_first[b] = -1;
    for (int l = 0; l < mctgen::kNumLayers; ++l) {
        const mctgen::LayerRecipe& r = mctgen::kLayerRecipe[l];
        if (r.kind == 0) seed_layer(S, seed, l, scratch);
        else if (blended_first[r.a] < 0) blended_first[r.a] = l;
    }
    for (int b = 0; b < mctgen::kNumBlended; ++b) seed_blended(S, seed, b, blended_first[b], scratch);
    seed_misc(S, seed, e.M);
    for (int i = 0; i < kCornerYTab; ++i) seed_yaxis(S, i);
    e.seed_value = seed;
    e.has_seed = true;
}

KeyCodec codec_for(const std::vector<ChunkRef>& ch, int xlo, int xhi, int zlo, int zhi) {
    return make_codec(ch.data(), static_cast<int>(ch.size()), 4, xlo, xhi, zlo, zhi);
}

// unique sorted keys and the map key -> unique index
void dedupe(const std::vector<uint64_t>& keys, std::vector<uint64_t>& uniq, std::vector<int>& map) {
    uniq = keys;
    std::sort(uniq.begin(), uniq.end());
    uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
    map.resize(keys.size());
    for (size_t i = 0; i < keys.size(); ++i) map[i] = static_cast<int>(std::lower_bound(uniq.begin(), uniq.end(), keys[i]) - uniq.begin());
}

// Mismatches (bit patterns) of the device's squeeze division against t / 24.0f over the
// floats of magnitude bits [lo, hi), both signs. Compiled twice: plain, and for CPUs with a
// fused multiply-add instruction, where __builtin_fmaf is that instruction instead of a call.
MCT_FI int64_t squeeze_div_range(uint32_t lo, uint32_t hi) {
    int64_t bad = 0;
    for (uint32_t u = lo; u < hi; ++u)
        for (uint32_t s = 0; s < 2; ++s) {
            const float t = u2f(u | (s << 31));
            bad += f2u(squeeze_div24(t)) != f2u(t / 24.0f) ? 1 : 0;
        }
    return bad;
}
int64_t squeeze_div_range_plain(uint32_t lo, uint32_t hi) { return squeeze_div_range(lo, hi); }
__attribute__((target("fma"))) int64_t squeeze_div_range_fma(uint32_t lo, uint32_t hi) { return squeeze_div_range(lo, hi); }

// java.lang.Math.min / max (float) and Mth.clamp(float) as the JDK writes them: a NaN first
// operand is returned, then negative zero is below positive zero, then the compare. The
// references for the select forms in jmath.h.
float java_minf(float a, float b) {
    if (a != a) return a;
    if (a == 0.0f && b == 0.0f && (f2u(b) >> 31)) return b;
    return a <= b ? a : b;
}
float java_maxf(float a, float b) {
    if (a != a) return a;
    if (a == 0.0f && b == 0.0f && (f2u(a) >> 31)) return b;
    return a >= b ? a : b;
}
float java_clampf(float v, float lo, float hi) { return v < lo ? lo : java_minf(v, hi); }

}  // namespace

extern "C" {

const char* mcte_last_error() { return g_err.c_str(); }

void* mcte_create() {
    auto e = new Emu;
    fill_model(*e);
    return e;
}

void mcte_destroy(void* p) { delete static_cast<Emu*>(p); }

void mcte_set_replay_all(void* p, int on) { static_cast<Emu*>(p)->replay_all = on != 0; }

// Seed tables as the device builds them: per layer {ox, oy, oz} and 256 perm bytes.
int mcte_seed_tables(void* p, int64_t seed, double* offsets, uint8_t* perms, int64_t* zoom_seed, uint64_t* positional6, int32_t* clay192) {
    Emu& e = *static_cast<Emu*>(p);
    build_seed(e, seed);
    const SeedDev& S = *e.seed;
    for (int l = 0; l < mctgen::kNumLayers; ++l) {
        if (offsets) { offsets[3 * l] = S.oct[l].ox; offsets[3 * l + 1] = S.oct[l].oy; offsets[3 * l + 2] = S.oct[l].oz; }
        if (perms)
            for (int i = 0; i < 256; ++i) perms[256 * l + i] = static_cast<uint8_t>(S.oct[l].q[i] & 0xFF);
    }
    if (zoom_seed) *zoom_seed = S.zoom_seed;
    if (positional6) {
        positional6[0] = S.random.lo; positional6[1] = S.random.hi;
        positional6[2] = S.aquifer.lo; positional6[3] = S.aquifer.hi;
        positional6[4] = S.ore.lo; positional6[5] = S.ore.hi;
    }
    if (clay192) for (int i = 0; i < 192; ++i) clay192[i] = S.clay[i];
    return mctgen::kNumLayers;
}

// Biome-only queries for n blocks (x, y, z triples) under one seed; -1 entries where the float
// search cannot serve the query.
int mcte_biome_query(void* p, int64_t seed, const int32_t* xyz, int n, int32_t* out) {
    try {
        Emu& e = *static_cast<Emu*>(p);
        build_seed(e, seed);
        for (int i = 0; i < n; ++i) out[i] = biome_query(*e.seed, e.M, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

// The column-cached zoom against zoom_pick on n columns (x, z pairs), every y: mismatches.
int64_t mcte_zoom_check(int64_t zoom_seed, const int32_t* xz, int n) {
    int64_t bad = 0;
    for (int i = 0; i < n; ++i) {
        ZoomColumn zc;
        zc.init(zoom_seed, xz[2 * i], xz[2 * i + 1]);
        for (int y = kMinY + kHeight - 1; y >= kMinY; --y) {
            int a, b, c, qa, qb, qc;
            zc.pick(y, a, b, c);
            zoom_pick(zoom_seed, xz[2 * i], y, xz[2 * i + 1], qa, qb, qc);
            bad += (a != qa || b != qb || c != qc) ? 1 : 0;
        }
    }
    return bad;
}

// Operation counts of stage i of the last mcte_generate on this thread: its name and
// {FP64, FP32, INT32, memory reads}; returns the number of stages, 0 before any call. All
// counts are zero unless the library was built with MCT_COUNT_OPS (mcte_counts_ops).
int mcte_stage_ops(int i, const char** name, uint64_t* ops) {
    const int n = static_cast<int>(g_stages.size());
    if (i >= 0 && i < n) {
        if (name) *name = g_stages[static_cast<size_t>(i)].name;
        if (ops) std::memcpy(ops, g_stages[static_cast<size_t>(i)].ops, sizeof g_stages[0].ops);
    }
    return n;
}

// The device's divisions by constants against the IEEE division, bit for bit:
//   out[0]      squeeze_div24 over every float in [-1, 1] (threads host threads);
//   out[1 + i]  div_known(x, d_i, RN(1 / d_i)) over every half-integer |x| <= 2^19, for the
//               aquifer divisors d = 25, 1.5, 2.5, 3, 10;
//   out[6 + i]  the uncorrected product x * RN(1 / d_i) over the same x, which the check must
//               see differ (its own sensitivity);
//   out[11]     the device's fused gradient terms gmad(g, v, c) = fma(g, v, c) against the
//               unfused g * v + c, for g in {-1, 0, 1}, over 2^24 pseudo-random (v, c) with
//               |v| < 2 and |c| < 4 (the dots' fractions and partial sums) and every pair of
//               +-0, +-2^-149, +-2^-126, +-1.
// Mismatch counts; all but out[6..10] are zero when the device's rewrites are exact.
int mcte_check_divisions(int threads, int64_t* out) {
    const uint32_t end = 0x3F800001u;   // magnitudes 0 .. 1.0f
    threads = std::max(1, std::min(threads, 64));
    const bool fma = __builtin_cpu_supports("fma");
    std::vector<int64_t> part(static_cast<size_t>(threads), 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            const uint32_t lo = static_cast<uint32_t>(static_cast<uint64_t>(end) * t / threads);
            const uint32_t hi = static_cast<uint32_t>(static_cast<uint64_t>(end) * (t + 1) / threads);
            part[static_cast<size_t>(t)] = fma ? squeeze_div_range_fma(lo, hi) : squeeze_div_range_plain(lo, hi);
        });
    for (auto& th : pool) th.join();
    out[0] = 0;
    for (int64_t v : part) out[0] += v;
    const double ds[5] = {25.0, 1.5, 2.5, 3.0, 10.0};
    for (int i = 0; i < 5; ++i) {
        const double d = ds[i], y = 1.0 / d;
        int64_t bad = 0, raw = 0;
        for (int64_t k = -(int64_t(1) << 20); k <= (int64_t(1) << 20); ++k) {
            const double x = static_cast<double>(k) * 0.5;
            uint64_t a, b, c;
            const double q = x / d, f = div_known(x, d, y), p = x * y;
            std::memcpy(&a, &q, 8);
            std::memcpy(&b, &f, 8);
            std::memcpy(&c, &p, 8);
            bad += a != b ? 1 : 0;
            raw += a != c ? 1 : 0;
        }
        out[1 + i] = bad;
        out[6 + i] = raw;
    }
    int64_t gbad = 0;
    const auto gcheck = [&gbad](float v, float c) {
        for (int g = -1; g <= 1; ++g) {
            const float gf = static_cast<float>(g);
            const float a = __builtin_fmaf(gf, v, c), b = gf * v + c;
            gbad += f2u(a) != f2u(b) ? 1 : 0;
        }
    };
    uint64_t st = 0;
    for (int i = 0; i < (1 << 24); ++i) {
        uint64_t r = (st += 0x9E3779B97F4A7C15ull);   // SplitMix64
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
        r ^= r >> 31;
        // sign, an exponent below 128 (magnitude < 2) or 129 (< 4), any mantissa
        const uint32_t v = (static_cast<uint32_t>(r) & 0x807FFFFFu) | ((static_cast<uint32_t>(r >> 23) % 128u) << 23);
        const uint32_t c = (static_cast<uint32_t>(r >> 32) & 0x807FFFFFu) | ((static_cast<uint32_t>(r >> 55) % 129u) << 23);
        gcheck(u2f(v), u2f(c));
    }
    const float sp[8] = {0.0f, -0.0f, 0x1p-149f, -0x1p-149f, 0x1p-126f, -0x1p-126f, 1.0f, -1.0f};
    for (float v : sp)
        for (float c : sp) gcheck(v, c);
    out[11] = gbad;
    return 0;
}

// jminf / jmaxf / clampf (jmath.h, select forms) against the JDK's branching definitions,
// bit for bit: out[0] every pair of 20 special floats (zeros, NaNs of both signs, infinities,
// subnormals, ones) for min and max, and every triple for clamp; out[1] 2^26 pseudo-random
// bit-pattern pairs for min and max; out[2] clamp of every float v for the bounds (-1, 1),
// (0, 4), (-0.0, -0.0), (0.0, -0.0) (threads host threads). Mismatch counts.
int mcte_check_minmax(int threads, int64_t* out) {
    const float sp[20] = {0.0f, -0.0f, 1.0f, -1.0f, u2f(0x7FC00000u), u2f(0xFFC00000u), u2f(0x7F800001u), u2f(0x7F800000u), u2f(0xFF800000u),
                          u2f(1u), u2f(0x80000001u), 0x1p-126f, -0x1p-126f, 4.0f, -64.0f, 320.0f, 0.5f, 3.5f, -3.5f, 0x1p127f};
    int64_t bad = 0;
    for (float a : sp)
        for (float b : sp) {
            bad += f2u(jminf(a, b)) != f2u(java_minf(a, b)) ? 1 : 0;
            bad += f2u(jmaxf(a, b)) != f2u(java_maxf(a, b)) ? 1 : 0;
            for (float c : sp) bad += f2u(clampf(a, b, c)) != f2u(java_clampf(a, b, c)) ? 1 : 0;
        }
    out[0] = bad;
    bad = 0;
    uint64_t st = 7;
    for (int i = 0; i < (1 << 26); ++i) {
        uint64_t r = (st += 0x9E3779B97F4A7C15ull);   // SplitMix64
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
        r ^= r >> 31;
        const float a = u2f(static_cast<uint32_t>(r)), b = u2f(static_cast<uint32_t>(r >> 32));
        bad += f2u(jminf(a, b)) != f2u(java_minf(a, b)) ? 1 : 0;
        bad += f2u(jmaxf(a, b)) != f2u(java_maxf(a, b)) ? 1 : 0;
    }
    out[1] = bad;
    const float bounds[4][2] = {{-1.0f, 1.0f}, {0.0f, 4.0f}, {-0.0f, -0.0f}, {0.0f, -0.0f}};
    threads = std::max(1, std::min(threads, 64));
    std::vector<int64_t> part(static_cast<size_t>(threads), 0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            const uint64_t lo = (uint64_t(1) << 32) * static_cast<uint64_t>(t) / static_cast<uint64_t>(threads);
            const uint64_t hi = (uint64_t(1) << 32) * static_cast<uint64_t>(t + 1) / static_cast<uint64_t>(threads);
            int64_t b = 0;
            for (uint64_t u = lo; u < hi; ++u) {
                const float v = u2f(static_cast<uint32_t>(u));
                for (const auto& bd : bounds) b += f2u(clampf(v, bd[0], bd[1])) != f2u(java_clampf(v, bd[0], bd[1])) ? 1 : 0;
            }
            part[static_cast<size_t>(t)] = b;
        });
    for (auto& th : pool) th.join();
    out[2] = 0;
    for (int64_t v : part) out[2] += v;
    return 0;
}

// The dense dedupe's key index (dense_index / dense_key, pipeline.h) over every key of random
// batches: n chunks (seeds from `nseeds` slots) around a few centres within +-spread chunks, for
// the column, grid and biome codecs: out[0] keys whose index is outside [0, dense_size), out[1]
// keys that dense_key does not give back, out[2] adjacent distinct sorted keys whose indexes are
// not increasing, out[3] batches checked.
int mcte_check_dense(int batches, int n, int nseeds, int spread, int64_t* out) {
    std::memset(out, 0, 4 * sizeof(int64_t));
    uint64_t st = 12345;
    const auto rnd = [&st](uint64_t m) {
        uint64_t r = (st += 0x9E3779B97F4A7C15ull);
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
        return (r ^ (r >> 31)) % m;
    };
    for (int b = 0; b < batches; ++b) {
        std::vector<ChunkRef> ch(static_cast<size_t>(n));
        const int centres = 1 + static_cast<int>(rnd(4));
        int cx[4], cz[4];
        for (int k = 0; k < centres; ++k) {
            cx[k] = static_cast<int>(rnd(2 * 117187 + 1)) - 117187;
            cz[k] = static_cast<int>(rnd(2 * 117187 + 1)) - 117187;
        }
        for (auto& c : ch) {
            const int k = static_cast<int>(rnd(static_cast<uint64_t>(centres)));
            c.x = cx[k] + static_cast<int>(rnd(2 * static_cast<uint64_t>(spread) + 1)) - spread;
            c.z = cz[k] + static_cast<int>(rnd(2 * static_cast<uint64_t>(spread) + 1)) - spread;
            c.seed = static_cast<int>(rnd(static_cast<uint64_t>(nseeds)));
            c.pad = 0;
        }
        const struct { int unit, xlo, xhi, zlo, zhi; } cs[3] = {{4, 0, 4, 0, 4}, {4, -16, 10, -8, 10}, {1, -1, 1, -1, 1}};
        for (const auto& c : cs) {
            const KeyCodec K = make_codec(ch.data(), n, c.unit, c.xlo, c.xhi, c.zlo, c.zhi);
            std::vector<uint64_t> keys;
            for (const auto& r : ch)
                for (int dx = c.xlo; dx <= c.xhi; ++dx)
                    for (int dz = c.zlo; dz <= c.zhi; ++dz)
                        for (int f = 0; f < 2; ++f) keys.push_back(make_key(K, r.seed, f, r.x * c.unit + dx, r.z * c.unit + dz));
            std::sort(keys.begin(), keys.end());
            keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
            const uint64_t D = dense_size(K);
            for (size_t i = 0; i < keys.size(); ++i) {
                const uint32_t d = dense_index(K, keys[i]);
                out[0] += d >= D ? 1 : 0;
                out[1] += dense_key(K, d) != keys[i] ? 1 : 0;
                if (i > 0) out[2] += d <= dense_index(K, keys[i - 1]) ? 1 : 0;
            }
        }
        ++out[3];
    }
    return 0;
}

int mcte_counts_ops() {
#if defined(MCT_COUNT_OPS)
    return 1;
#else
    return 0;
#endif
}

// Full pipeline for one seed and n chunks. Outputs (nullable): terrain and aquifer fill u8 per
// block, biomes u8 per quart, density f32 per block (canonical layouts, chunks back to back).
// stats[0] = chunks redone in exact order after the speculative surface pass, stats[1] = their
// columns walked again, stats[2] = those of the chunks whose blocks the replay changed. With replay_all set
// (mcte_set_replay_all) every column of a redone chunk is refilled and walked again, in the
// x-major order, instead of only those whose steep test the live heights change.
int mcte_generate(void* p, int64_t seed, const int32_t* xz, int n, uint8_t* terrain, uint8_t* biomes, uint8_t* aquifer, float* density, int64_t* stats) {
    try {
        Emu& e = *static_cast<Emu*>(p);
        build_seed(e, seed);
        stage_begin();
        const SeedDev* seeds = e.seed.get();
        const ModelDev& M = e.M;
        std::vector<ChunkRef> ch(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) ch[i] = {xz[2 * i], xz[2 * i + 1], 0, 0};
        // biome chunks: union of the 3x3 neighbourhoods
        std::unordered_map<uint64_t, int> bindex;
        std::vector<ChunkRef> bch;
        std::vector<int> nb(static_cast<size_t>(n) * 9);
        for (int i = 0; i < n; ++i)
            for (int dz = -1; dz <= 1; ++dz)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int bx = ch[i].x + dx, bz = ch[i].z + dz;
                    const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(bx)) << 32) | static_cast<uint32_t>(bz);
                    auto it = bindex.find(key);
                    if (it == bindex.end()) {
                        it = bindex.emplace(key, static_cast<int>(bch.size())).first;
                        bch.push_back({bx, bz, 0, 0});
                    }
                    nb[static_cast<size_t>(i) * 9 + (dz + 1) * 3 + dx + 1] = it->second;
                }
        const int nbk = static_cast<int>(bch.size());
        std::vector<float> bcol(static_cast<size_t>(nbk) * 16 * (mctgen::kNumClimateColumns > 0 ? mctgen::kNumClimateColumns : 1));
        std::vector<int64_t> targets(kOneVarying ? 1 : static_cast<size_t>(nbk) * kQPer * 7), mind(static_cast<size_t>(nbk) * kQPer);
        std::vector<int64_t> tcol(kOneVarying ? static_cast<size_t>(nbk) * 16 * 7 : 1);
        std::vector<int32_t> lstar(static_cast<size_t>(nbk) * kQPer);
        std::vector<uint8_t> tie(static_cast<size_t>(nbk) * kQPer), bb(static_cast<size_t>(nbk) * kQPer);
        for (int i = 0; i < nbk * 16; ++i) biome_column(seeds, bch.data(), bcol.data(), i);
        stage_mark("biome_columns");
        std::vector<int64_t> tvary(static_cast<size_t>(nbk) * kQPer);
        const Targets T{targets.data(), tcol.data(), tvary.data()};
        for (int i = 0; i < nbk * kQPer; ++i) biome_target(seeds, bch.data(), bcol.data(), T, i);
        const int ntop = M.treef ? std::min(mctgen::kTreeNodes, 384) : 0;
        const int G = (M.treef && kOneVarying) ? M.col_groups[kVaryCoord] : 0;
        if (G > 0) {
            std::vector<GroupRes> gres(static_cast<size_t>(nbk) * 16 * G);
            for (int i = 0; i < nbk * 16 * G; ++i) biome_group_search(M, M.treef, ntop, T, lstar.data(), mind.data(), gres.data(), G, i);
            for (int i = 0; i < nbk * kQPer; ++i) biome_quart_final(M, tvary.data(), gres.data(), lstar.data(), mind.data(), G, i);
        } else {
            for (int i = 0; i < nbk * 16; ++i) biome_search_column(M, M.treef, ntop, T, lstar.data(), mind.data(), i);
        }
        stage_mark("biome_quarts");
        for (int i = 0; i < nbk * kQPer; ++i) biome_search_wide(M, T, lstar.data(), mind.data(), i);
        for (int i = 0; i < nbk * kQPer; ++i) biome_tie(M, T, lstar.data(), mind.data(), tie.data(), i);
        stage_mark("biome_ties");
        for (int b = 0; b < nbk; ++b) biome_chain_seq(M, T, lstar.data(), mind.data(), bb.data(), b);
        stage_mark("biome_chain");
        // corner columns and corners
        std::vector<uint64_t> ckeys(static_cast<size_t>(n) * kCols), ucols;
        std::vector<int> colmap;
        const KeyCodec KC = codec_for(ch, 0, 4, 0, 4);
        for (int i = 0; i < n * kCols; ++i) column_key(KC, ch.data(), ckeys.data(), i);
        dedupe(ckeys, ucols, colmap);
        const int U = static_cast<int>(ucols.size());
        std::vector<float> colv(static_cast<size_t>(U) * (mctgen::kNumColumns > 0 ? mctgen::kNumColumns : 1));
        std::vector<float> corners(static_cast<size_t>(U) * kCh * kRows);
        stage_mark("dedupe_columns");
        for (int u = 0; u < U; ++u) column_values(KC, seeds, ucols.data(), colv.data(), u);
        stage_mark("columns");
        std::vector<mct::AxisXZ> tab(static_cast<size_t>(mctgen::kCornerSites > 0 ? mctgen::kCornerSites : 1));
        for (int u = 0; u < U; ++u) {
            for (int k = 0; k < mctgen::kCornerSites; ++k) tab[k] = corner_site_axes(KC, seeds, ucols.data(), u, k);
            for (int row = 0; row < kRows; ++row) corner_values(KC, seeds, ucols.data(), colv.data(), corners.data(), u, row, tab.data());
        }
        stage_mark("corners");
        // surface-level grid and aquifer
        std::vector<uint64_t> gkeys(static_cast<size_t>(n) * kGridKeys), ug;
        std::vector<int> gmap;
        const KeyCodec KG = codec_for(ch, -16, 10, -8, 10);
        for (int i = 0; i < n * kGridKeys; ++i) grid_key(KG, ch.data(), gkeys.data(), i);
        dedupe(gkeys, ug, gmap);
        stage_mark("dedupe_grid");
        std::vector<float> gv(ug.size());
        for (size_t u = 0; u < ug.size(); ++u) fts_value(KG, seeds, ug.data(), gv.data(), static_cast<int>(u));
        stage_mark("fts");
        std::vector<int32_t> grid(static_cast<size_t>(n) * kGridCells);
        std::vector<float> prelim(static_cast<size_t>(n) * 4);
        for (int i = 0; i < n * kGridKeys; ++i) grid_scatter(gmap.data(), gv.data(), grid.data(), prelim.data(), i);
        std::vector<int> skip(static_cast<size_t>(n));
        for (int c = 0; c < n; ++c) aquifer_skip(grid.data(), skip.data(), c);
        stage_mark("grid_scatter+aquifer_skip");
        std::vector<int32_t> cloc(static_cast<size_t>(n) * kAqCells * 4), cfl(static_cast<size_t>(n) * kAqCells * 2);
        for (int i = 0; i < n * kAqCells; ++i) aquifer_cell(seeds, ch.data(), grid.data(), cloc.data(), cfl.data(), i);
        stage_mark("aquifer_cells");
        // fill
        std::vector<uint8_t> blocks(static_cast<size_t>(n) * kPer);
        FillArgs F{seeds, ch.data(), corners.data(), colmap.data(), cloc.data(), cfl.data(), skip.data(), blocks.data()};
        for (int c = 0; c < n; ++c)
            for (int t = 0; t < 256 * kCellsY; ++t) fill_cell(F, c, t, density ? density + static_cast<size_t>(c) * kPer : nullptr);
        stage_mark("fill");
        if (aquifer) std::memcpy(aquifer, blocks.data(), blocks.size());
        // surface
        std::vector<int> h0(static_cast<size_t>(n) * 256), h1(h0.size()), h2(h0.size()), hf(h0.size());
        for (int c = 0; c < n; ++c)
            for (int col = 0; col < 256; ++col) h0[c * 256 + col] = column_height(M, blocks.data() + static_cast<size_t>(c) * kPer, col >> 4, col & 15);
        stage_mark("heights");
        std::vector<float> colrec(static_cast<size_t>(n) * 256 * kColF);
        for (int i = 0; i < n * 256; ++i) surface_column(seeds, M, ch.data(), colrec.data(), i);
        stage_mark("surface_columns");
        SurfArgs A{seeds, &M, ch.data(), blocks.data(), nb.data(), bb.data(), corners.data(), colmap.data(), prelim.data(), colrec.data()};
        for (int c = 0; c < n; ++c)
            for (int col = 0; col < 256; ++col) surface_pass1(A, h0.data(), h1.data(), h2.data(), c, col);
        stage_mark("surface1");
        for (int c = 0; c < n; ++c)
            for (int col = 0; col < 256; ++col) surface_pass2<false>(A, h0.data(), h1.data(), h2.data(), hf.data(), c, col);
        stage_mark("surface2+ore_veins");
        int64_t redone = 0, recols = 0, changed = 0;
        for (int c = 0; c < n; ++c) {
            bool ok = true;
            for (int col = 0; col < 256; ++col) ok &= hf[c * 256 + col] == h2[c * 256 + col];
            if (ok) continue;
            ++redone;
            // the device's selective replay in x-major order (a linear order of its
            // anti-diagonals): a column whose steep test the live heights leave as the
            // speculative pass had it keeps its result, any other is refilled and walked again
            const int* h0c = h0.data() + static_cast<size_t>(c) * 256;
            const int* h1c = h1.data() + static_cast<size_t>(c) * 256;
            const int* h2c = h2.data() + static_cast<size_t>(c) * 256;
            const std::vector<uint8_t> speculative(blocks.begin() + static_cast<ptrdiff_t>(c) * kPer, blocks.begin() + static_cast<ptrdiff_t>(c + 1) * kPer);
            int hm[256];
            for (int col = 0; col < 256; ++col) hm[col] = h0[c * 256 + col];
            for (int col = 0; col < 256; ++col) {
                if (!e.replay_all && !surface_replay_needed(hm, h0c, h1c, h2c, col)) {
                    hm[col] = hf[c * 256 + col];
                    continue;
                }
                for (int cy = 0; cy < kCellsY; ++cy) fill_cell(F, c, (col >> 4) + (col & 15) * 16 + cy * 256, nullptr);
                surface_exact_column(A, hm, c, col);
                ++recols;
            }
            changed += std::memcmp(speculative.data(), blocks.data() + static_cast<size_t>(c) * kPer, kPer) != 0;
        }
        stage_mark("surface_check+exact");
        if (stats) {
            stats[0] = redone;
            stats[1] = recols;
            stats[2] = changed;
        }
        if (terrain) std::memcpy(terrain, blocks.data(), blocks.size());
        if (biomes)
            for (int c = 0; c < n; ++c) std::memcpy(biomes + static_cast<size_t>(c) * kQPer, bb.data() + static_cast<size_t>(nb[c * 9 + 4]) * kQPer, kQPer);
        return 0;
    } catch (const std::exception& ex) {
        g_err = ex.what();
        return 1;
    }
}

}  // extern "C"
