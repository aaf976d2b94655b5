// C driver over the reference cpu/ port, compiled from the reference's own sources by
// oracle/Makefile into oracle/_ref/libmcref.so. Test infrastructure only: it lets the
// tests and the CPU baseline call the unmodified reference implementation
// (cpu/src/terrain.cpp:825-885 cpu_generate_terrain / cpu_terrain_from_biomes,
// cpu/src/volume_eval.cpp:111-509 sample / point, cpu/src/seed_tables.cpp:159-184).
#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "mccuda/biome.h"
#include "mccuda/ir.h"
#include "mccuda/rng.h"
#include "mccuda/seed_tables.h"
#include "mccuda/terrain.h"
#include "mccuda/terrain_model.h"
#include "mccuda/volume_eval.h"

using namespace mcc;

namespace {

struct Handle {
    IR ir;
    TerrainModel tm;
    int64_t seed = 0;
    std::unique_ptr<TerrainSeed> ts;
    std::string err;
};

thread_local std::string g_err;

TerrainSeed& seed_of(Handle* h, int64_t seed) {
    if (!h->ts || h->seed != seed) {
        h->ts = std::make_unique<TerrainSeed>(h->ir, h->tm, seed);
        h->seed = seed;
    }
    return *h->ts;
}

template <typename F>
void parallel_for(int threads, int count, F&& f) {
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&, t] {
            int i;
            while ((i = next.fetch_add(1)) < count) f(t, i);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_create(const char* ir_path, const char* terrain_path) {
    try {
        auto h = new Handle;
        h->ir = load_ir(ir_path);
        h->tm = load_terrain_model(terrain_path);
        return h;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void ref_destroy(void* p) { delete static_cast<Handle*>(p); }

// Full TERRAIN of independent chunks (each with its own 3x3 biome neighbourhood).
// Outputs per chunk: biomes u8[16*qh], aquifer u16[16*16*h], terrain u16[16*16*h]
// (any may be null), density f32[16*16*h] (nullable).
int ref_generate_terrain(void* p, int64_t seed, const int32_t* xz, int n, uint8_t* biomes, uint16_t* aquifer, uint16_t* terrain, float* density) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        const size_t per = static_cast<size_t>(16 * 16 * h->ir.height), qper = static_cast<size_t>(16 * (h->ir.height / 4));
        for (int i = 0; i < n; ++i) {
            ChunkTerrain c;
            cpu_generate_terrain(ts, xz[2 * i], xz[2 * i + 1], c);
            if (biomes) std::memcpy(biomes + i * qper, c.biomes.data(), qper);
            if (aquifer) std::memcpy(aquifer + i * per, c.aquifer.data(), per * 2);
            if (terrain) std::memcpy(terrain + i * per, c.terrain.data(), per * 2);
            if (density) std::memcpy(density + i * per, c.density.data(), per * 4);
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Core configuration of independent chunks (cpu/src/volume_eval.cpp:676-690 cpu_generate_chunk):
// final density f32 and core fill u8 per block (either may be null).
int ref_generate_core(void* p, int64_t seed, const int32_t* xz, int n, float* density, uint8_t* fill) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        const size_t per = static_cast<size_t>(16 * 16 * h->ir.height);
        VolumeEvaluator ev(h->ir, ts.tables);
        for (int i = 0; i < n; ++i)
            cpu_generate_chunk(h->ir, ev, xz[2 * i], xz[2 * i + 1], density ? density + i * per : nullptr, fill ? fill + i * per : nullptr);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// A w x d region at chunk (x0, z0), generated the way the reference bench does
// (bench/bench_main.cpp:253-304): biomes of the region plus a one-chunk border once,
// then fill + surface per chunk, on `threads` std::threads. Outputs row-major by chunk.
int ref_generate_region(void* p, int64_t seed, int x0, int z0, int w, int d, int threads, uint8_t* biomes, uint16_t* terrain) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        const IR& ir = h->ir;
        const size_t per = static_cast<size_t>(16 * 16 * ir.height);
        const int qper = 16 * (ir.height / 4);
        const int bw = w + 2, bd = d + 2;
        std::vector<uint8_t> bb(static_cast<size_t>(bw) * bd * qper);
        std::vector<std::unique_ptr<VolumeEvaluator>> evals(static_cast<size_t>(threads));
        parallel_for(threads, bw * bd, [&](int t, int i) {
            if (!evals[t]) evals[t] = std::make_unique<VolumeEvaluator>(ir, ts.tables);
            cpu_chunk_biomes(ts, *evals[t], x0 - 1 + i % bw, z0 - 1 + i / bw, bb.data() + static_cast<size_t>(i) * qper);
        });
        parallel_for(threads, w * d, [&](int t, int i) {
            int x = i % w, z = i / w;
            const uint8_t* nb[9];
            for (int dz = 0; dz < 3; ++dz)
                for (int dx = 0; dx < 3; ++dx) nb[dz * 3 + dx] = bb.data() + static_cast<size_t>((z + dz) * bw + x + dx) * qper;
            if (!evals[t]) evals[t] = std::make_unique<VolumeEvaluator>(ir, ts.tables);
            ChunkTerrain c;
            cpu_terrain_from_biomes(ts, *evals[t], x0 + x, z0 + z, nb, c);
            if (terrain) std::memcpy(terrain + static_cast<size_t>(i) * per, c.terrain.data(), per * 2);
            if (biomes) std::memcpy(biomes + static_cast<size_t>(i) * qper, c.biomes.data(), static_cast<size_t>(qper));
        });
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Biomes of one chunk as doCreateBiomes stores them (fresh R-tree candidate).
int ref_chunk_biomes(void* p, int64_t seed, int cx, int cz, uint8_t* out) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        VolumeEvaluator ev(h->ir, ts.tables);
        cpu_chunk_biomes(ts, ev, cx, cz, out);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Volume-path sample of a named root over Volume{sx,sy,sz, minx,miny,minz, stx,sty,stz}
// (output index y + (x + z*sx)*sy), with fresh caches.
int ref_sample(void* p, int64_t seed, const char* root, const int32_t* vol, float* out) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        VolumeEvaluator ev(h->ir, ts.tables);
        Volume v{vol[0], vol[1], vol[2], vol[3], vol[4], vol[5], vol[6], vol[7], vol[8]};
        ev.sample(h->ir.root(root), v, out);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Volume-path sample of any IR node id (debug bisection), fresh caches.
int ref_sample_id(void* p, int64_t seed, int node, const int32_t* vol, float* out) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        VolumeEvaluator ev(h->ir, ts.tables);
        Volume v{vol[0], vol[1], vol[2], vol[3], vol[4], vol[5], vol[6], vol[7], vol[8]};
        ev.sample(node, v, out);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Point-path value of any IR node id at n positions.
int ref_point_id(void* p, int64_t seed, int node, const int32_t* xyz, int n, float* out) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        VolumeEvaluator ev(h->ir, ts.tables);
        for (int i = 0; i < n; ++i) out[i] = ev.point(node, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Point-path values of a named root at n positions (xyz triples).
int ref_point(void* p, int64_t seed, const char* root, const int32_t* xyz, int n, float* out) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        VolumeEvaluator ev(h->ir, ts.tables);
        const int id = h->ir.root(root);
        for (int i = 0; i < n; ++i) out[i] = ev.point(id, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Per-seed noise tables flattened: for each layer (in SeedTables order) 6 doubles
// {frequency, amplitude, ox, oy, oz, smear} and 256 perm bytes. Returns the layer count
// (call with null buffers to size them).
int ref_seed_layers(void* p, int64_t seed, double* params, uint8_t* perms) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        const SeedTables& t = ts.tables;
        if (params) {
            for (size_t i = 0; i < t.layers.size(); ++i) {
                const StackLayer& L = t.layers[i];
                const PerlinOctave& o = t.octaves[static_cast<size_t>(L.octave)];
                double* q = params + 6 * i;
                q[0] = L.frequency; q[1] = static_cast<double>(L.amplitude); q[2] = o.ox; q[3] = o.oy; q[4] = o.oz; q[5] = o.smear;
                std::memcpy(perms + 256 * i, o.perm, 256);
            }
        }
        return static_cast<int>(t.layers.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Layer range [first, first + count) of a noise key in ref_seed_layers order.
int ref_noise_ref(void* p, int64_t seed, const char* key, int* first, int* count) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        const StackRef& s = ts.tables.noise.at(key);
        *first = s.first_layer;
        *count = s.count;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Per-seed scalars: zoom seed, positional factories (random, aquifer, ore) and clay bands.
int ref_seed_misc(void* p, int64_t seed, int64_t* zoom_seed, uint64_t* positional6, int32_t* clay192) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, seed);
        *zoom_seed = ts.zoom_seed;
        positional6[0] = ts.random.lo(); positional6[1] = ts.random.hi();
        positional6[2] = ts.aquifer_random.lo(); positional6[3] = ts.aquifer_random.hi();
        positional6[4] = ts.ore_random.lo(); positional6[5] = ts.ore_random.hi();
        for (int i = 0; i < 192; ++i) clay192[i] = ts.clay_bands[i];
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// R-tree search (cpu/src/biome.cpp:224-228) for n targets, each with its own candidate leaf.
int ref_tree_search(void* p, const int64_t* targets, const int32_t* candidates, int n, int32_t* out) {
    try {
        Handle* h = static_cast<Handle*>(p);
        TerrainSeed& ts = seed_of(h, h->ts ? h->seed : 0);
        for (int i = 0; i < n; ++i) out[i] = ts.tree.search(targets + 7 * i, candidates[i]);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// The 64-bit chunk digest of the library's k_chunk_digest: the sum (mod 2^64) over the 8-byte
// little-endian words w_i of mix(w_i + (i + 1) * 0x9E3779B97F4A7C15), finalised as
// mix(sum ^ byte count), with mix the SplitMix64 finaliser.
static uint64_t digest_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t chunk_digest(const uint8_t* p, size_t bytes) {
    uint64_t h = 0;
    for (size_t i = 0; i < bytes / 8; ++i) {
        uint64_t w;
        std::memcpy(&w, p + 8 * i, 8);
        h += digest_mix(w + 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(i + 1));
    }
    return digest_mix(h ^ static_cast<uint64_t>(bytes));
}

// Digests of the full TERRAIN step (cpu_generate_terrain, each chunk with its own 3x3 biome
// neighbourhood) of n chunks, chunk i under seeds[i], on `threads` std::threads: terrain as u8
// palette indices (98,304 B) and biomes (1,536 B), each with chunk_digest. Work is handed out
// in runs of `run` consecutive chunks; each thread keeps the TerrainSeed of the last seed it
// met, so a list sorted by seed builds each seed's tables about once per run.
int ref_digest_terrain(void* p, const int64_t* seeds, const int32_t* xz, int64_t n, int threads, int run, uint64_t* dig_terrain, uint64_t* dig_biomes) {
    try {
        Handle* h = static_cast<Handle*>(p);
        const size_t per = static_cast<size_t>(16 * 16 * h->ir.height), qper = static_cast<size_t>(16 * (h->ir.height / 4));
        std::atomic<int64_t> next{0};
        std::atomic<int> failed{0};
        std::vector<std::string> errs(static_cast<size_t>(threads));
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t)
            pool.emplace_back([&, t] {
                try {
                    std::unique_ptr<TerrainSeed> ts;
                    int64_t have = 0;
                    std::vector<uint8_t> t8(per);
                    int64_t i0;
                    while (!failed.load() && (i0 = next.fetch_add(run)) < n) {
                        for (int64_t i = i0; i < std::min(n, i0 + run); ++i) {
                            if (!ts || have != seeds[i]) {
                                ts = std::make_unique<TerrainSeed>(h->ir, h->tm, seeds[i]);
                                have = seeds[i];
                            }
                            ChunkTerrain c;
                            cpu_generate_terrain(*ts, xz[2 * i], xz[2 * i + 1], c);
                            for (size_t k = 0; k < per; ++k) t8[k] = static_cast<uint8_t>(c.terrain[k]);
                            dig_terrain[i] = chunk_digest(t8.data(), per);
                            dig_biomes[i] = chunk_digest(c.biomes.data(), qper);
                        }
                    }
                } catch (const std::exception& e) {
                    errs[static_cast<size_t>(t)] = e.what();
                    failed = 1;
                }
            });
        for (auto& th : pool) th.join();
        for (auto& e : errs)
            if (!e.empty()) {
                g_err = e;
                return 1;
            }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// Answers of n biome-only block queries on `threads` threads: the biome the TERRAIN step stores
// for the block's zoomed quart (quart y clamped to the world height), i.e. the chunk biomes
// (cpu_chunk_biomes, fresh R-tree candidate per chunk) at the quart BiomeManager's zoom picks.
int ref_query_biomes(void* p, const int64_t* seeds, const int32_t* xyz, int64_t n, int threads, int run, uint8_t* out) {
    try {
        Handle* h = static_cast<Handle*>(p);
        const int qh = h->ir.height / 4, qmin = h->ir.min_y >> 2;
        std::atomic<int64_t> next{0};
        std::atomic<int> failed{0};
        std::vector<std::string> errs(static_cast<size_t>(threads));
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t)
            pool.emplace_back([&, t] {
                try {
                    std::unique_ptr<TerrainSeed> ts;
                    std::unique_ptr<VolumeEvaluator> ev;
                    int64_t have = 0;
                    std::vector<uint8_t> b(static_cast<size_t>(16 * qh));
                    int64_t i0;
                    while (!failed.load() && (i0 = next.fetch_add(run)) < n) {
                        for (int64_t i = i0; i < std::min(n, i0 + run); ++i) {
                            if (!ts || have != seeds[i]) {
                                ev.reset();
                                ts = std::make_unique<TerrainSeed>(h->ir, h->tm, seeds[i]);
                                ev = std::make_unique<VolumeEvaluator>(h->ir, ts->tables);
                                have = seeds[i];
                            }
                            int qx = 0, qy = 0, qz = 0;
                            zoomed_biome(ts->zoom_seed, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], [&](int a, int bb, int c) { qx = a; qy = bb; qz = c; return 0; });
                            cpu_chunk_biomes(*ts, *ev, qx >> 2, qz >> 2, b.data());
                            const int cqy = std::min(std::max(qy, qmin), qmin + qh - 1) - qmin;
                            out[i] = b[static_cast<size_t>((cqy * 4 + (qz & 3)) * 4 + (qx & 3))];
                        }
                    }
                } catch (const std::exception& e) {
                    errs[static_cast<size_t>(t)] = e.what();
                    failed = 1;
                }
            });
        for (auto& th : pool) th.join();
        for (auto& e : errs)
            if (!e.empty()) {
                g_err = e;
                return 1;
            }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// BiomeManager zoom: the quart (qx, qy, qz) picked for each block.
int ref_zoom(int64_t zoom_seed, const int32_t* xyz, int n, int32_t* out_q) {
    for (int i = 0; i < n; ++i) {
        int qx = 0, qy = 0, qz = 0;
        zoomed_biome(zoom_seed, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], [&](int a, int b, int c) { qx = a; qy = b; qz = c; return 0; });
        out_q[3 * i] = qx; out_q[3 * i + 1] = qy; out_q[3 * i + 2] = qz;
    }
    return 0;
}

}  // extern "C"
