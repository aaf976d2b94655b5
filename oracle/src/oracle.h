// Oracle: a CPU restatement of the reference's TERRAIN semantics, used only as the checker of
// the HIP library in tests, in __graft_entry__.smoke() and as bench.py's CPU baseline. It is
// never linked into or called by the product. Single-threaded per generator; independent
// generators may run on separate threads.
//
// What it restates (reference files, bonquifo/minecraft-terrain-cuda):
//   Java arithmetic            cpu/include/mccuda/javamath.h:61-181, df_ops.h:14-80
//   random sources, MD5        cpu/src/rng.cpp:19-210
//   gradient noise             cpu/include/mccuda/noise_eval.h:21-304
//   per-seed noise tables      cpu/src/seed_tables.cpp:23-184
//   IR / terrain model formats cpu/src/ir.cpp:63-157, cpu/src/terrain_model.cpp:111-169
//   density interpreter        cpu/src/volume_eval.cpp:38-690
//   climate R-tree, zoom       cpu/src/biome.cpp:14-289
//   biomes, aquifer, surface   cpu/src/terrain.cpp:46-885
//
// How each part is pinned (the tests under tests/ hold the checks): RFC 1321 / FIPS 180-4
// vectors for MD5 / SHA-256, the reference's in-source known answers (tests/test_l0.cpp:36-85),
// closed forms and invariants of the noise, brute force for the R-tree, and golden vectors
// written by the reference cpu/ port compiled from its sources (tests/golden/).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

namespace orc {

// ---- Java arithmetic -------------------------------------------------------------------------
inline uint32_t fb(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint64_t db(double d) { uint64_t u; std::memcpy(&u, &d, 8); return u; }
inline double bd(uint64_t u) { double d; std::memcpy(&d, &u, 8); return d; }

inline int32_t to_int(double v) {                    // (int) v
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return static_cast<int32_t>(v);
}
inline int32_t floor_i(double v) { return to_int(std::floor(v)); }   // Mth.floor
inline int32_t wadd(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b)); }
inline int32_t wmul(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) * static_cast<uint32_t>(b)); }
inline int32_t floor_div(int32_t a, int32_t b) {
    const int32_t q = a / b;
    return (q * b != a && ((a < 0) != (b < 0))) ? q - 1 : q;
}
inline int32_t floor_mod(int32_t a, int32_t b) { return a - floor_div(a, b) * b; }
float java_min(float a, float b);
float java_max(float a, float b);
double java_min(double a, double b);
double java_max(double a, double b);
inline float fclamp(float v, float lo, float hi) { return v < lo ? lo : java_min(v, hi); }
inline double dclamp(double v, double lo, double hi) { return v < lo ? lo : java_min(v, hi); }
inline float lerp(float t, float a, float b) { return a + t * (b - a); }
float lerp3(float a1, float a2, float a3, float x000, float x100, float x010, float x110, float x001, float x101, float x011, float x111);
float smoothstep(float x);

// ---- random sources ---------------------------------------------------------------------------
void md5(const std::string& s, uint8_t out[16]);
void sha256(const uint8_t* msg, size_t len, uint8_t out[32]);
int64_t zoom_seed_of(int64_t world_seed);            // BiomeManager.obfuscateSeed
uint64_t mix13(uint64_t z);                          // RandomSupport.mixStafford13

struct Xrand {                                       // XoroshiroRandomSource
    uint64_t s0, s1;
    Xrand(uint64_t lo, uint64_t hi);
    static Xrand from_seed(int64_t seed);            // upgradeSeedTo128bit
    uint64_t next();
    int32_t next_int(int32_t bound);
    double next_double();
    float next_float();
    bool next_bool() { return (next() & 1) != 0; }
};

struct Factory {                                     // XoroshiroPositionalRandomFactory
    uint64_t lo = 0, hi = 0;
    static Factory fork(Xrand& r);
    Xrand hash(const std::string& name) const;       // fromHashOf
    Xrand at(int32_t x, int32_t y, int32_t z) const;
};
int64_t block_seed(int32_t x, int32_t y, int32_t z);   // Mth.getSeed

struct Lcg {                                         // LegacyRandomSource
    uint64_t s;
    explicit Lcg(int64_t seed) : s((static_cast<uint64_t>(seed) ^ 0x5DEECE66Dull) & ((1ull << 48) - 1)) {}
    int32_t next(int bits);
    int32_t next_int(int32_t bound);
    double next_double();
};

// ---- gradient noise -----------------------------------------------------------------------------
struct Octave {
    double xo, yo, zo, smear;
    uint8_t perm[256];
};
struct Layer {
    double freq;
    float amp;
    int octave;
};
struct Stack {
    int first = 0, count = 0;
    bool smeared = false;
};

float noise_get(const Octave& o, double x, double y, double z);                     // PerlinNoise.get
float noise_get_smeared(const Octave& o, double x, double y, double z);             // SmearedPerlinNoise.get
float noise_add_volume(const Octave& o, int bx, int by, int bz, double xz, double ys);          // PerlinNoise.addToVolume (one sample)
float noise_add_volume_smeared(const Octave& o, int bx, int by, int bz, double xz, double ys);  // SmearedPerlinNoise.addToVolume
double wrap_coord(double x);                         // GradientNoise.wrap

// ---- JSON (parsing the two input files) ------------------------------------------------------
struct Json {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    bool b = false;
    std::string s;                                   // string, or the number's text
    std::vector<Json> a;
    std::vector<std::pair<std::string, Json>> o;
    const Json& operator[](const std::string& k) const;
    const Json& operator[](size_t i) const { return a.at(i); }
    bool has(const std::string& k) const;
    int64_t i64() const;
    uint32_t u32() const { return static_cast<uint32_t>(i64()); }
    float f32bits() const { return bf(u32()); }
    double hexd() const;                             // "0x..." raw bits
    size_t size() const { return kind == Arr ? a.size() : o.size(); }
};
Json parse_json_file(const std::string& path);

// ---- IR ---------------------------------------------------------------------------------------
enum class T {
    Cache, Const, Simple, Gradient, Noise, ShiftA, ShiftB, Shift, Blended, Binary, Unary, Clamp, Lerp, RangeChoice, IntervalSelect,
    Spline, Interp, Slice, BlendDensity, Fts
};

struct SplineN {
    bool is_const = false;
    float k = 0;
    int coord = -1;
    std::vector<float> loc, der;
    std::vector<SplineN> val;
};

struct Node {
    T t = T::Const;
    float lo = 0, hi = 0;                            // value range
    std::string op, kind, axis, tiling, noise;
    int in = -1, l = -1, r = -1, alpha = -1, first = -1, second = -1, when_in = -1, when_out = -1;
    int sx = -1, sy = -1, sz = -1;
    bool zx = true, zy = true, zz = true;            // shift_*_zero
    float v = 0, fmin = 0, fmax = 0, fv = 0, tv = 0;
    int from = 0, to = 0, coord = 0, cell_xz = 0, cell_y = 0;
    double xz = 0, ys = 0, xzf = 0, yf = 0, smear = 0;
    std::vector<float> th;
    std::vector<int> fns;
    int density = -1, upper = -1, lower = 0, cell_h = 0;
    std::shared_ptr<SplineN> spline;
};

struct NoiseParams {
    double base_amp = 1;
    int base_octave = 0, count = 1;
    std::string normalize;
    std::vector<double> mods;
};

struct Ir {
    int min_y = 0, height = 0, sea = 0;
    std::vector<Node> nodes;
    std::map<std::string, int> roots;
    std::map<std::string, NoiseParams> noises;
};
Ir load_ir(const std::string& path);

// ---- terrain model ----------------------------------------------------------------------------
struct Cond {
    std::string t;
    uint64_t mask[2] = {0, 0};
    std::string noise, random_name;
    double mn = 0, mx = 0;
    bool is3d = false, add_stone = false, add_surface = false, ceiling = false;
    int true_at = 0, false_at = 0, anchor = 0, mult = 0, offset = 0, secondary = 0;
    int child = -1;
};
struct Rule {
    std::string t;
    int state = -1, cond = -1, then = -1, ore = -1, raw = -1, filler = -1;
    std::vector<int> seq;
    float raw_chance = 0;
    std::string density, richness, gap;
};
struct Box {
    int64_t mn[7], mx[7];
    int biome;
};
struct Model {
    int sea = 0;
    std::vector<std::string> palette;
    std::vector<uint8_t> is_air, is_water, is_lava;
    int default_block = 0, default_fluid = 0, air = -1, water = -1, lava = -1, packed_ice = -1, snow = -1;
    int terracotta[7];
    std::vector<std::string> biome_names;
    std::vector<float> biome_temp;
    std::vector<uint8_t> biome_frozen;
    std::vector<Box> params;
    std::vector<Cond> conds;
    std::vector<Rule> rules;
    int root = -1;
    int eroded_badlands = -1, frozen_ocean = -1, deep_frozen_ocean = -1;
};
Model load_model(const std::string& path);

// ---- per-seed tables ---------------------------------------------------------------------------
struct Tables {
    std::vector<Octave> oct;
    std::vector<Layer> layers;
    std::map<std::string, Stack> noise;
    std::map<int, Stack[3]> blended;                 // per node: min_limit, max_limit, main
};
Tables make_tables(const Ir& ir, int64_t seed);
float stack_get(const Tables& t, const Stack& s, double x, double y, double z);
double stream_sum(const std::vector<double>& v);     // DoubleStream.sum

// ---- climate R-tree ------------------------------------------------------------------------------
class RTree {
public:
    explicit RTree(const std::vector<Box>& params);
    int search(const int64_t t[7], int candidate_leaf) const;   // leaf index
    int64_t leaf_dist(int leaf, const int64_t t[7]) const;
    int brute_first_min(const int64_t t[7], int64_t* d) const;  // first minimal leaf in DFS order (checker)

private:
    struct N {
        int64_t mn[7], mx[7];
        std::vector<int> kids;
        int leaf = -1;
    };
    int build(std::vector<int> items);
    int add(const std::vector<int>& kids);
    int64_t dist(int n, const int64_t t[7]) const;
    int walk(int n, const int64_t t[7], int cand) const;
    std::vector<N> n_;
    std::vector<int> leaf_node_;
    int root_ = -1;
};
int64_t quantize(float v);                          // Climate.quantizeCoord
void zoom_quart(int64_t zseed, int x, int y, int z, int& qx, int& qy, int& qz);   // BiomeManager.getBiome

// ---- density-function interpreter ----------------------------------------------------------------
struct Vol {
    int sx, sy, sz, x0, y0, z0, dx = 1, dy = 1, dz = 1;
    int n() const { return sx * sy * sz; }
    int at(int x, int y, int z) const { return y + (x + z * sx) * sy; }
    bool operator==(const Vol& o) const;
};

class Density {
public:
    Density(const Ir& ir, const Tables& t);
    void sample(int node, const Vol& v, float* out);     // volume path (sampleVolume)
    float point(int node, int x, int y, int z);          // point path (sampleValue)
    void reset() { cache_.clear(); }

private:
    void fill_noise(const Stack& s, const Vol& v, double xz, double ys, float amp, float* acc);
    void interpolated(const Node& n, const Vol& v, float* out);
    void interp_blocks(const Node& n, const Vol& v, float* out);
    void slice(const Node& n, const Vol& v, float* out);
    float spline_vol(const SplineN& s, int i, std::map<int, std::vector<float>>& in, const Vol& v);
    float spline_pt(const SplineN& s, int x, int y, int z);
    float surface_probe(const Node& n, int x, int z, float upper);
    const Ir& ir_;
    const Tables& t_;
    std::unordered_map<int, std::pair<Vol, std::vector<float>>> cache_;
};

// ---- TERRAIN ----------------------------------------------------------------------------------------
struct World {
    const Ir& ir;
    const Model& m;
    int64_t seed;
    Tables tables;
    RTree tree;
    int64_t zseed;
    Factory random, aquifer, ore;
    std::map<std::string, Factory> gradient;
    int clay[192];
    uint8_t simplex[5][256];                         // temperature, frozen x3, biome info
    World(const Ir& ir, const Model& m, int64_t seed);
};

void chunk_biomes(World& w, Density& d, int cx, int cz, uint8_t* out);
// Fill + surface of one chunk from the biomes of its 3x3 neighbourhood (nb[(dz+1)*3 + dx+1]).
void chunk_terrain(World& w, Density& d, int cx, int cz, const uint8_t* const nb[9], uint16_t* aquifer, uint16_t* terrain, float* density);
float biome_temperature(const World& w, int biome, int x, int y, int z);

}  // namespace orc
