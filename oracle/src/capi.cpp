// C entry points of the oracle for the tests (ctypes). Test infrastructure only.
#include <algorithm>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <string>
#include <thread>

#include "oracle.h"

using namespace orc;

namespace {
thread_local std::string g_err;

struct Handle {
    Ir ir;
    Model m;
    std::map<int64_t, std::unique_ptr<World>> worlds;
    World& world(int64_t seed) {
        auto it = worlds.find(seed);
        if (it == worlds.end()) it = worlds.emplace(seed, std::make_unique<World>(ir, m, seed)).first;
        return *it->second;
    }
};

template <class F>
int guard(F f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

uint64_t chunk_key(int cx, int cz) { return (static_cast<uint64_t>(static_cast<uint32_t>(cz)) << 32) | static_cast<uint32_t>(cx); }
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_create(const char* ir_path, const char* terrain_path) {
    Handle* h = nullptr;
    guard([&] {
        auto p = std::make_unique<Handle>();
        p->ir = load_ir(ir_path);
        p->m = load_model(terrain_path);
        h = p.release();
    });
    return h;
}

void orc_destroy(void* h) { delete static_cast<Handle*>(h); }

// Chunks (cx, cz) pairs; per chunk 16*(h/4) biomes, 16*16*h aquifer/terrain states and
// densities (index (y*16 + z)*16 + x); any output may be null. Chunks are spread over `threads`.
int orc_generate(void* hp, int64_t seed, const int32_t* xz, int n, int threads, uint8_t* biomes, uint16_t* aquifer, uint16_t* terrain, float* density) {
    return guard([&] {
        Handle& h = *static_cast<Handle*>(hp);
        World& w = h.world(seed);
        const int qn = 16 * (w.ir.height / 4);
        const size_t per = static_cast<size_t>(256) * static_cast<size_t>(w.ir.height);
        // biomes of every chunk in the 3x3 neighbourhoods, each computed once
        std::map<uint64_t, std::vector<uint8_t>> bio;
        for (int i = 0; i < n; ++i)
            for (int dz = -1; dz <= 1; ++dz)
                for (int dx = -1; dx <= 1; ++dx) bio[chunk_key(xz[2 * i] + dx, xz[2 * i + 1] + dz)];
        std::vector<std::pair<uint64_t, std::vector<uint8_t>*>> todo;
        for (auto& kv : bio) todo.emplace_back(kv.first, &kv.second);
        const int T = std::max(1, std::min(threads, 64));
        std::vector<std::string> errs(static_cast<size_t>(T));
        auto par = [&](size_t count, auto body) {
            std::vector<std::thread> pool;
            for (int t = 0; t < T; ++t)
                pool.emplace_back([&, t] {
                    try {
                        Density d(w.ir, w.tables);
                        for (size_t k = static_cast<size_t>(t); k < count; k += static_cast<size_t>(T)) body(d, k);
                    } catch (const std::exception& e) {
                        errs[static_cast<size_t>(t)] = e.what();
                    }
                });
            for (auto& th : pool) th.join();
            for (auto& e : errs)
                if (!e.empty()) throw std::runtime_error(e);
        };
        par(todo.size(), [&](Density& d, size_t k) {
            const uint64_t key = todo[k].first;
            todo[k].second->resize(static_cast<size_t>(qn));
            chunk_biomes(w, d, static_cast<int32_t>(static_cast<uint32_t>(key)), static_cast<int32_t>(static_cast<uint32_t>(key >> 32)), todo[k].second->data());
        });
        par(static_cast<size_t>(n), [&](Density& d, size_t i) {
            const int cx = xz[2 * i], cz = xz[2 * i + 1];
            const uint8_t* nb[9];
            for (int dz = -1; dz <= 1; ++dz)
                for (int dx = -1; dx <= 1; ++dx) nb[(dz + 1) * 3 + dx + 1] = bio.at(chunk_key(cx + dx, cz + dz)).data();
            if (biomes) std::memcpy(biomes + i * static_cast<size_t>(qn), nb[4], static_cast<size_t>(qn));
            chunk_terrain(w, d, cx, cz, nb, aquifer ? aquifer + i * per : nullptr, terrain ? terrain + i * per : nullptr, density ? density + i * per : nullptr);
        });
    });
}

// Volume path of one node: vol = {sx, sy, sz, x0, y0, z0, dx, dy, dz}; out in (z, x, y) order.
int orc_sample(void* hp, int64_t seed, int node, const int32_t* vol, float* out) {
    return guard([&] {
        Handle& h = *static_cast<Handle*>(hp);
        World& w = h.world(seed);
        Density d(w.ir, w.tables);
        const Vol v{vol[0], vol[1], vol[2], vol[3], vol[4], vol[5], vol[6], vol[7], vol[8]};
        d.sample(node, v, out);
    });
}

// Point path of one node at n positions (x, y, z triples).
int orc_point(void* hp, int64_t seed, int node, const int32_t* xyz, int n, float* out) {
    return guard([&] {
        Handle& h = *static_cast<Handle*>(hp);
        World& w = h.world(seed);
        Density d(w.ir, w.tables);
        for (int i = 0; i < n; ++i) out[i] = d.point(node, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    });
}

int orc_root(void* hp, const char* name) {
    const Handle& h = *static_cast<Handle*>(hp);
    auto it = h.ir.roots.find(name);
    return it == h.ir.roots.end() ? -1 : it->second;
}

int orc_node_count(void* hp) { return static_cast<int>(static_cast<Handle*>(hp)->ir.nodes.size()); }

// Climate search: n targets of 7 quantized coordinates; out_leaf = the search result with the
// candidate chain (previous result, starting at none); out_brute = the first leaf at minimal
// distance over all leaves in tree order, out_dist its distance, out_cdist the search result's.
int orc_tree(void* hp, const int64_t* t, int n, int chain, int32_t* out_leaf, int32_t* out_brute, int64_t* out_dist, int64_t* out_cdist) {
    return guard([&] {
        Handle& h = *static_cast<Handle*>(hp);
        const RTree tree(h.m.params);
        int last = -1;
        for (int i = 0; i < n; ++i) {
            const int64_t* ti = t + 7 * i;
            last = tree.search(ti, chain ? last : -1);
            out_leaf[i] = last;
            int64_t d = 0;
            out_brute[i] = tree.brute_first_min(ti, &d);
            out_dist[i] = d;
            out_cdist[i] = tree.leaf_dist(last, ti);
        }
    });
}

int orc_leaf_biome(void* hp, int leaf) { return static_cast<Handle*>(hp)->m.params.at(static_cast<size_t>(leaf)).biome; }

int64_t orc_quantize(float v) { return quantize(v); }

void orc_zoom(int64_t zseed, const int32_t* xyz, int n, int32_t* out) {
    for (int i = 0; i < n; ++i) zoom_quart(zseed, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], out[3 * i], out[3 * i + 1], out[3 * i + 2]);
}
int64_t orc_zoom_seed(int64_t seed) { return zoom_seed_of(seed); }

// ---- known-answer hooks for the arithmetic and the random sources ----
void orc_md5(const char* s, uint8_t* out) { md5(s, out); }
void orc_sha256(const uint8_t* msg, size_t len, uint8_t* out) { sha256(msg, len, out); }
void orc_xoro(int64_t seed, int n, uint64_t* out) {
    Xrand r = Xrand::from_seed(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next();
}
void orc_xoro_ints(int64_t seed, int bound, int n, int32_t* out) {
    Xrand r = Xrand::from_seed(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next_int(bound);
}
void orc_xoro_doubles(int64_t seed, int n, double* out) {
    Xrand r = Xrand::from_seed(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next_double();
}
void orc_lcg_ints(int64_t seed, int bound, int n, int32_t* out) {
    Lcg r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next_int(bound);
}
void orc_lcg_next(int64_t seed, int bits, int n, int32_t* out) {
    Lcg r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next(bits);
}
void orc_lcg_doubles(int64_t seed, int n, double* out) {
    Lcg r(seed);
    for (int i = 0; i < n; ++i) out[i] = r.next_double();
}
uint64_t orc_mix13(uint64_t z) { return mix13(z); }
int64_t orc_block_seed(int32_t x, int32_t y, int32_t z) { return block_seed(x, y, z); }
// fromHashOf(name).at(x, y, z).nextLong() under the world seed's positional factory
uint64_t orc_positional(int64_t seed, const char* name, int32_t x, int32_t y, int32_t z) {
    Xrand base = Xrand::from_seed(seed);
    const Factory f = Factory::fork(base);
    Xrand r = f.hash(name);
    const Factory g = Factory::fork(r);
    return g.at(x, y, z).next();
}
int32_t orc_to_int(double v) { return to_int(v); }
int32_t orc_floor(double v) { return floor_i(v); }
float orc_fmin(float a, float b) { return java_min(a, b); }
float orc_fmax(float a, float b) { return java_max(a, b); }
double orc_wrap(double v) { return wrap_coord(v); }
double orc_stream_sum(const double* v, int n) { return stream_sum(std::vector<double>(v, v + n)); }

// One octave with an explicit permutation and offsets: the noise at (x, y, z) (plain and smeared).
float orc_octave(const uint8_t* perm, double xo, double yo, double zo, double smear, double x, double y, double z, int smeared) {
    Octave o{};
    o.xo = xo;
    o.yo = yo;
    o.zo = zo;
    o.smear = smear;
    std::memcpy(o.perm, perm, 256);
    return smeared ? noise_get_smeared(o, x, y, z) : noise_get(o, x, y, z);
}

float orc_biome_temperature(void* hp, int64_t seed, int biome, int x, int y, int z) {
    Handle& h = *static_cast<Handle*>(hp);
    return biome_temperature(h.world(seed), biome, x, y, z);
}

// Per-seed tables: the octave count, then for octave k its offsets and permutation.
int orc_octaves(void* hp, int64_t seed, int k, double* offs, uint8_t* perm) {
    Handle& h = *static_cast<Handle*>(hp);
    const World& w = h.world(seed);
    if (k < 0 || k >= static_cast<int>(w.tables.oct.size())) return static_cast<int>(w.tables.oct.size());
    const Octave& o = w.tables.oct[static_cast<size_t>(k)];
    offs[0] = o.xo;
    offs[1] = o.yo;
    offs[2] = o.zo;
    offs[3] = o.smear;
    std::memcpy(perm, o.perm, 256);
    return static_cast<int>(w.tables.oct.size());
}

// The layers of noise `key` under `seed`, in stack order: 6 doubles per layer {frequency,
// amplitude, ox, oy, oz, smear} and 256 permutation bytes. Returns the layer count (-1: no key).
int orc_noise_layers(void* hp, int64_t seed, const char* key, double* params, uint8_t* perms, int cap) {
    Handle& h = *static_cast<Handle*>(hp);
    const World& w = h.world(seed);
    auto it = w.tables.noise.find(key);
    if (it == w.tables.noise.end()) return -1;
    const Stack& s = it->second;
    for (int i = 0; i < s.count && i < cap; ++i) {
        const Layer& L = w.tables.layers[static_cast<size_t>(s.first + i)];
        const Octave& o = w.tables.oct[static_cast<size_t>(L.octave)];
        double* q = params + 6 * i;
        q[0] = L.freq; q[1] = static_cast<double>(L.amp); q[2] = o.xo; q[3] = o.yo; q[4] = o.zo; q[5] = o.smear;
        std::memcpy(perms + 256 * i, o.perm, 256);
    }
    return s.count;
}

void orc_clay(void* hp, int64_t seed, int32_t* out) {
    Handle& h = *static_cast<Handle*>(hp);
    std::memcpy(out, h.world(seed).clay, sizeof(int) * 192);
}

}  // extern "C"
