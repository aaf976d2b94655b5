// Drives libmct.so through the C++ class (include/mct_terrain_gpu.hpp) the way a caller of
// the reference's FullTerrainGenerator does: set_seed, generate with timing, tie statistics.
// Usage: generator_example <ir.json> <terrain.json> <seed> <out.bin> cx cz [cx cz ...]
// Writes terrain (98,304 B per chunk) then biomes (1,536 B per chunk) to out.bin.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mct_terrain_gpu.hpp"

int main(int argc, char** argv) {
    if (argc < 7 || (argc - 5) % 2 != 0) {
        std::fprintf(stderr, "usage: %s ir.json terrain.json seed out.bin cx cz [cx cz ...]\n", argv[0]);
        return 2;
    }
    try {
        std::vector<mct::ChunkXZ> chunks;
        for (int i = 5; i + 1 < argc; i += 2) chunks.push_back({std::atoi(argv[i]), std::atoi(argv[i + 1])});
        const int n = static_cast<int>(chunks.size());
        mct::FullTerrainGenerator gen(argv[1], argv[2], /*max_batch_chunks=*/2, /*pipeline_slots=*/2);
        gen.set_seed(std::strtoll(argv[3], nullptr, 10));
        uint8_t* terrain = mct::alloc_pinned(static_cast<size_t>(n) * 98304);
        std::vector<uint8_t> biomes(static_cast<size_t>(n) * 1536);
        mct::TerrainTiming t;
        gen.generate(chunks.data(), n, terrain, biomes.data(), nullptr, &t);
        const mct::BiomeTieStats st = gen.biome_tie_stats();
        const std::vector<double> done = gen.last_batch_done_ms();
        std::FILE* f = std::fopen(argv[4], "wb");
        if (!f) throw std::runtime_error("cannot write output");
        std::fwrite(terrain, 1, static_cast<size_t>(n) * 98304, f);
        std::fwrite(biomes.data(), 1, biomes.size(), f);
        std::fclose(f);
        mct::free_pinned(terrain);
        std::printf("chunks %d batches %zu kernel_ms %.3f quarts %lld\n", n, done.size(), t.kernel_ms(), static_cast<long long>(st.quarts));
        // the error path: a bad batch size throws std::runtime_error with the library's message
        try {
            mct::FullTerrainGenerator bad(argv[1], argv[2], 0, 1);
            return 1;
        } catch (const std::runtime_error&) {
        }
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
