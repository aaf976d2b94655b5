// C++ face of libmct.so: the reference's generator class, FullTerrainGenerator
// (cuda/include/mccuda_cuda/terrain_gpu.h:13-63 in bonquifo/minecraft-terrain-cuda), over the
// C ABI in mct.h. Header-only; link with -lmct.
//
// Same names, argument meaning and error behaviour as the reference class: outputs in the
// reference's layouts, a std::runtime_error carrying the library's message on any failure,
// one host thread per generator. The one difference is the constructor's model arguments:
// the reference takes parsed IR / TerrainModel objects held by reference; this one takes the
// paths of the same mccuda-ir/1 and mccuda-terrain/1 files (the library compiles the model
// in, checks the files against it, and owns everything it needs).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "mct.h"

namespace mct {

struct ChunkXZ {          // generator.h:14-16
    int32_t x, z;
};

struct TerrainTiming {    // terrain_gpu.h:19-22
    double h2d_ms = 0, biome_ms = 0, noise_ms = 0, aquifer_ms = 0, fill_ms = 0, surface_ms = 0, d2h_ms = 0, wall_ms = 0;
    double kernel_ms() const { return biome_ms + noise_ms + aquifer_ms + fill_ms + surface_ms; }
};

struct BiomeTieStats {    // terrain_gpu.h:27-29
    int64_t quarts, overrides, biome_changes;
};

inline void check(int rc) {
    if (rc != MCT_OK) throw std::runtime_error(std::string("mct: ") + mct_last_error());
}

class FullTerrainGenerator {
public:
    FullTerrainGenerator(const std::string& ir_json, const std::string& terrain_json, int max_batch_chunks = 256, int pipeline_slots = 2,
                         int seed_capacity = 64) {
        check(mct_create(ir_json.c_str(), terrain_json.c_str(), max_batch_chunks, pipeline_slots, seed_capacity, &g_));
    }
    ~FullTerrainGenerator() { mct_destroy(g_); }
    FullTerrainGenerator(const FullTerrainGenerator&) = delete;
    FullTerrainGenerator& operator=(const FullTerrainGenerator&) = delete;

    void set_seed(int64_t seed) { check(mct_set_seed(g_, seed)); }

    // terrain_out: u8 palette index per block, ((y - minY) * 16 + z) * 16 + x, chunks back to
    // back; biomes_out (u8 per quart, (qy * 4 + qz) * 4 + qx) and aquifer_out may be null.
    void generate(const ChunkXZ* chunks, int count, uint8_t* terrain_out, uint8_t* biomes_out, uint8_t* aquifer_out, TerrainTiming* timing = nullptr) {
        mct_timing t{};
        check(mct_generate(g_, reinterpret_cast<const int32_t*>(chunks), count, terrain_out, biomes_out, aquifer_out, timing ? &t : nullptr));
        if (timing) *timing = TerrainTiming{t.h2d_ms, t.biome_ms, t.noise_ms, t.aquifer_ms, t.fill_ms, t.surface_ms, t.d2h_ms, t.wall_ms};
    }

    // Many seeds in one call (chunk i under seeds[i]); no counterpart in the reference class.
    void generate_multi(const int64_t* seeds, const ChunkXZ* chunks, int count, uint8_t* terrain_out, uint8_t* biomes_out, uint8_t* aquifer_out,
                        TerrainTiming* timing = nullptr) {
        mct_timing t{};
        check(mct_generate_multi(g_, seeds, reinterpret_cast<const int32_t*>(chunks), count, terrain_out, biomes_out, aquifer_out, timing ? &t : nullptr));
        if (timing) *timing = TerrainTiming{t.h2d_ms, t.biome_ms, t.noise_ms, t.aquifer_ms, t.fill_ms, t.surface_ms, t.d2h_ms, t.wall_ms};
    }

    BiomeTieStats biome_tie_stats() {
        int64_t s[4];
        check(mct_biome_tie_stats(g_, s));
        return BiomeTieStats{s[0], s[1], s[2]};
    }

    std::vector<double> last_batch_done_ms() const {
        int n = 0;
        check(mct_last_batch_done_ms(g_, nullptr, 0, &n));
        std::vector<double> v(static_cast<size_t>(n));
        if (n > 0) check(mct_last_batch_done_ms(g_, v.data(), n, &n));
        return v;
    }

    // Biome stored for each block xyz[3i..3i+2] under seeds[i], without generating chunks;
    // no counterpart in the reference class.
    void biome_query(const int64_t* seeds, const int32_t* xyz, int count, uint8_t* biomes_out) {
        check(mct_biome_query(g_, seeds, xyz, count, biomes_out));
    }

    mct_generator* handle() const { return g_; }

private:
    mct_generator* g_ = nullptr;
};

// Core configuration, the reference's TerrainGenerator (cuda/include/mccuda_cuda/generator.h:30-45):
// final density and the core fill (0 air, 1 stone, 2 water, 3 lava) per block, no biomes,
// aquifers or material rules. set_seed takes the world seed (the reference takes its SeedTables;
// the tables are built on the device here).
class TerrainGenerator {
public:
    TerrainGenerator(const std::string& ir_json, const std::string& terrain_json, int max_batch_chunks = 1024) {
        check(mct_create(ir_json.c_str(), terrain_json.c_str(), max_batch_chunks, 2, 0, &g_));
    }
    ~TerrainGenerator() { mct_destroy(g_); }
    TerrainGenerator(const TerrainGenerator&) = delete;
    TerrainGenerator& operator=(const TerrainGenerator&) = delete;

    void set_seed(int64_t seed) { seed_ = seed; }

    // density_out (f32) may be null; outputs in the canonical layout, chunks back to back.
    void generate(const ChunkXZ* chunks, int count, float* density_out, uint8_t* fill_out, TerrainTiming* timing = nullptr) {
        std::vector<int64_t> seeds(static_cast<size_t>(count), seed_);
        mct_timing t{};
        check(mct_generate_core(g_, seeds.data(), reinterpret_cast<const int32_t*>(chunks), count, density_out, fill_out, timing ? &t : nullptr));
        if (timing) *timing = TerrainTiming{t.h2d_ms, t.biome_ms, t.noise_ms, t.aquifer_ms, t.fill_ms, t.surface_ms, t.d2h_ms, t.wall_ms};
    }

private:
    mct_generator* g_ = nullptr;
    int64_t seed_ = 0;
};

struct ZoomSelfTest {     // terrain_gpu.h:56-58 (field names kept)
    int64_t checked, mismatches, fallbacks;   // fallbacks: near ties (a second candidate within 8e-6), resolved in FP64
};

// The surface pass's column-cached zoom against the plain FP64 zoom on about `count` blocks.
inline ZoomSelfTest zoom_selftest(int64_t world_seed, int64_t count) {
    int64_t out[3];
    check(mct_zoom_selftest(world_seed, count, out));
    return ZoomSelfTest{out[0], out[1], out[2]};
}

inline uint8_t* alloc_pinned(size_t bytes) {    // terrain_gpu.h:61
    void* p = mct_alloc_pinned(bytes);
    if (!p) throw std::runtime_error(std::string("mct: ") + mct_last_error());
    return static_cast<uint8_t*>(p);
}
inline void free_pinned(void* p) { mct_free_pinned(p); }

}  // namespace mct
