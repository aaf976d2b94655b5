// libmct.so: the gfx950 TERRAIN-step generator behind the C ABI of include/mct.h.
//
//   kernels.h   the HIP kernels (per-item bodies in pipeline.h, generated evaluators in gen/)
//   dedupe.h    device radix-sort dedupe of position keys
//   engine.h    batches as captured graphs on lanes, resident seed tables, copies back
//   server.h    the serving loop (continuous batching over an HBM result arena)
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off -fno-fast-math
//        -fno-gpu-flush-denormals-to-zero -fhip-fp32-correctly-rounded-divide-sqrt
// (Java semantics need unfused, correctly rounded IEEE arithmetic with denormals).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mct.h"
#include "engine.h"
#include "server.h"

using namespace mct;

// The digest buffer of mct_chunk_digests_noalloc: a file-scope device array (one per GPU, live
// until the process ends). Host code uses the address hipGetSymbolAddress returns for the
// current device, never the array's name (a host-side stand-in, not a device address).
static __device__ __align__(256) uint64_t scratch[MCT_DIGESTS_MAX];

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}

template <typename F>
int guard(F&& f) {
    try {
        f();
        return MCT_OK;
    } catch (const Error& e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc&) {
        return fail(MCT_ECAPACITY, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(MCT_EINVAL, e.what());
    }
}

uint64_t file_fnv1a(const char* path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error(MCT_EIO, std::string("cannot open ") + path);
    uint64_t h = 0xcbf29ce484222325ull;
    char buf[1 << 16];
    while (in) {
        in.read(buf, sizeof buf);
        const std::streamsize got = in.gcount();
        for (std::streamsize i = 0; i < got; ++i) {
            h ^= static_cast<uint8_t>(buf[i]);
            h *= 0x100000001b3ull;
        }
    }
    return h;
}

void check_model_files(const char* ir_json, const char* terrain_json) {
    if (!ir_json || !terrain_json) throw Error(MCT_EINVAL, "null path");
    if (file_fnv1a(ir_json) != mctgen::kIrHash) throw Error(MCT_EMODEL, "the IR file differs from the one this library was generated from (rebuild)");
    if (file_fnv1a(terrain_json) != mctgen::kTerrainHash)
        throw Error(MCT_EMODEL, "the terrain model file differs from the one this library was generated from (rebuild)");
}

// Chunk coordinates the pipeline takes: the world's chunk range with room for the position
// keys' neighbourhoods (|x|, |z| < 2^28).
constexpr int32_t kChunkLimit = 1 << 28;
void check_chunks(const int32_t* xz, int count) {
    for (int i = 0; i < 2 * count; ++i)
        if (xz[i] <= -kChunkLimit || xz[i] >= kChunkLimit)
            throw Error(MCT_EINVAL, "chunk " + std::to_string(i / 2) + ": coordinate " + std::to_string(xz[i]) + " outside +-2^28");
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[kMaxDevices] = {};

uint64_t* scratch_on_current_device() {
    int dev = 0;
    hip_check(hipGetDevice(&dev), "get device");
    if (dev < 0 || dev >= kMaxDevices) throw Error(MCT_ECAPACITY, "device index beyond the library's table");
    std::lock_guard<std::mutex> l(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "digest buffer address");
        g_scratch_ptr[dev] = static_cast<uint64_t*>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

struct mct_generator {
    std::unique_ptr<Engine> e;
    int64_t current_seed = 0;
};

struct mct_server {
    std::unique_ptr<Server> s;
};

extern "C" {

const char* mct_last_error(void) { return g_err.c_str(); }

int mct_create(const char* ir_json, const char* terrain_json, int max_batch_chunks, int pipeline_slots, int seed_capacity, mct_generator** out) {
    return guard([&] {
        if (!out) throw Error(MCT_EINVAL, "out is null");
        *out = nullptr;
        if (!ir_json || !terrain_json) throw Error(MCT_EINVAL, "null path");
        if (max_batch_chunks < 1 || max_batch_chunks > (1 << 16)) throw Error(MCT_EINVAL, "max_batch_chunks must be 1..65536");
        if (pipeline_slots < 1 || pipeline_slots > 4) throw Error(MCT_EINVAL, "pipeline_slots must be 1..4");
        if (seed_capacity == 0) seed_capacity = 64;
        if (seed_capacity < 1 || seed_capacity > 4096) throw Error(MCT_EINVAL, "seed_capacity must be 1..4096");
        check_model_files(ir_json, terrain_json);
        auto h = std::make_unique<mct_generator>();
        h->e = std::make_unique<Engine>(max_batch_chunks, pipeline_slots, seed_capacity);
        *out = h.release();
    });
}

int mct_destroy(mct_generator* g) {
    return guard([&] { delete g; });
}

static constexpr int64_t kQueryRange = int64_t{1} << 30;

int mct_biome_query(mct_generator* g, const int64_t* seeds, const int32_t* xyz, int count, uint8_t* biomes) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!seeds || !xyz || !biomes))) throw Error(MCT_EINVAL, "bad arguments");
        for (int i = 0; i < count; ++i)
            if (std::abs(static_cast<int64_t>(xyz[3 * i])) > kQueryRange || std::abs(static_cast<int64_t>(xyz[3 * i + 2])) > kQueryRange)
                throw Error(MCT_EINVAL, "query " + std::to_string(i) + ": |x| or |z| beyond 2^30");
        g->e->biome_query(seeds, xyz, count, biomes);
    });
}

int mct_zoom_selftest(int64_t world_seed, int64_t count, int64_t out[3]) {
    return guard([&] {
        if (!out || count < 0) throw Error(MCT_EINVAL, "bad arguments");
        const int64_t zoom_seed = static_cast<int64_t>(sha256_long_le(world_seed));
        unsigned long long* d = nullptr;
        hip_check(hipMalloc(&d, sizeof(unsigned long long) * 3), "malloc");
        try {
            hip_check(hipMemset(d, 0, sizeof(unsigned long long) * 3), "memset");
            const int64_t columns = (count + kHeight - 1) / kHeight;
            for (int64_t c0 = 0; c0 < columns; c0 += (1 << 24)) {
                const int nc = static_cast<int>(std::min<int64_t>(1 << 24, columns - c0));
                k_zoom_selftest<<<blocks_for(static_cast<size_t>(nc), 256), 256>>>(zoom_seed, 0x5A4F4F4Dull + static_cast<uint64_t>(c0), nc, d);
                hip_check(hipGetLastError(), "zoom selftest kernel");
            }
            unsigned long long h[3];
            hip_check(hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost), "download");
            out[0] = static_cast<int64_t>(h[0]);
            out[1] = static_cast<int64_t>(h[1]);
            out[2] = static_cast<int64_t>(h[2]);
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        hip_check(hipFree(d), "free");
    });
}

int mct_server_create(const char* ir_json, const char* terrain_json, int max_batch_chunks, int seed_capacity, size_t hbm_budget_bytes, mct_server** out) {
    return guard([&] {
        if (!out) throw Error(MCT_EINVAL, "out is null");
        *out = nullptr;
        if (max_batch_chunks < 1 || max_batch_chunks > (1 << 16)) throw Error(MCT_EINVAL, "max_batch_chunks out of range");
        if (seed_capacity == 0) seed_capacity = 64;
        if (seed_capacity < 1 || seed_capacity > 4096) throw Error(MCT_EINVAL, "seed_capacity must be 1..4096");
        check_model_files(ir_json, terrain_json);
        auto h = std::make_unique<mct_server>();
        h->s = std::make_unique<Server>(max_batch_chunks, seed_capacity, hbm_budget_bytes);
        *out = h.release();
    });
}

int mct_server_submit(mct_server* s, int64_t seed, const int32_t* chunk_xz, int count, uint8_t* terrain, uint8_t* biomes, int64_t* request_id) {
    return guard([&] {
        if (!s || count < 1 || !chunk_xz || !terrain || !request_id) throw Error(MCT_EINVAL, "bad arguments");
        check_chunks(chunk_xz, count);
        *request_id = s->s->submit(seed, chunk_xz, count, terrain, biomes);
    });
}

int mct_server_poll(mct_server* s, int64_t* ids, double* latency_ms, int cap, int* n) {
    return guard([&] {
        if (!s || !n || cap < 0 || (cap > 0 && !ids)) throw Error(MCT_EINVAL, "bad arguments");
        *n = s->s->poll(ids, latency_ms, cap);
    });
}

int mct_server_drain(mct_server* s) {
    return guard([&] {
        if (!s) throw Error(MCT_EINVAL, "null server");
        s->s->drain();
    });
}

int mct_server_stats(mct_server* s, int64_t out[8]) {
    return guard([&] {
        if (!s || !out) throw Error(MCT_EINVAL, "bad arguments");
        s->s->stats(out);
    });
}

int mct_server_destroy(mct_server* s) {
    return guard([&] { delete s; });
}

int mct_set_seed(mct_generator* g, int64_t seed) {
    return guard([&] {
        if (!g) throw Error(MCT_EINVAL, "null generator");
        g->e->finish_all(nullptr);
        g->current_seed = seed;
    });
}

int mct_generate(mct_generator* g, const int32_t* chunk_xz, int count, uint8_t* terrain, uint8_t* biomes, uint8_t* aquifer, mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!chunk_xz || !terrain))) throw Error(MCT_EINVAL, "bad arguments");
        check_chunks(chunk_xz, count);
        std::vector<int64_t> seeds(static_cast<size_t>(count), g->current_seed);
        BatchOut o;
        o.terrain = terrain;
        o.biomes = biomes;
        o.aquifer = aquifer;
        g->e->generate(seeds.data(), chunk_xz, count, o, timing);
    });
}

int mct_generate_multi(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, uint8_t* terrain, uint8_t* biomes, uint8_t* aquifer,
                       mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!seeds || !chunk_xz || !terrain))) throw Error(MCT_EINVAL, "bad arguments");
        check_chunks(chunk_xz, count);
        BatchOut o;
        o.terrain = terrain;
        o.biomes = biomes;
        o.aquifer = aquifer;
        g->e->generate(seeds, chunk_xz, count, o, timing);
    });
}

int mct_generate_device(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, uint8_t* d_terrain, uint8_t* d_biomes,
                        mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!seeds || !chunk_xz || !d_terrain))) throw Error(MCT_EINVAL, "bad arguments");
        check_chunks(chunk_xz, count);
        BatchOut o;
        o.terrain = d_terrain;
        o.biomes = d_biomes;
        o.device = true;
        g->e->generate(seeds, chunk_xz, count, o, timing);
    });
}

int mct_generate_core(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, float* density, uint8_t* fill, mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!seeds || !chunk_xz || (!density && !fill)))) throw Error(MCT_EINVAL, "bad arguments");
        check_chunks(chunk_xz, count);
        BatchOut o;
        o.terrain = fill;
        o.density = density;
        g->e->generate(seeds, chunk_xz, count, o, timing, true);
    });
}

int mct_generate_core_device(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, float* d_density, uint8_t* d_fill,
                             mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!seeds || !chunk_xz || (!d_density && !d_fill)))) throw Error(MCT_EINVAL, "bad arguments");
        check_chunks(chunk_xz, count);
        BatchOut o;
        o.terrain = d_fill;
        o.density = d_density;
        o.device = true;
        g->e->generate(seeds, chunk_xz, count, o, timing, true);
    });
}

int mct_autotune(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, mct_tuning* out) {
    return guard([&] {
        if (!g || count < 1 || !seeds || !chunk_xz) throw Error(MCT_EINVAL, "bad arguments");
        check_chunks(chunk_xz, count);
        g->e->autotune(seeds, chunk_xz, count, out);
    });
}

int mct_get_tuning(mct_generator* g, mct_tuning* out) {
    return guard([&] {
        if (!g || !out) throw Error(MCT_EINVAL, "bad arguments");
        g->e->get_tuning(out);
    });
}

int mct_set_tuning(mct_generator* g, const mct_tuning* t) {
    return guard([&] {
        if (!g || !t) throw Error(MCT_EINVAL, "bad arguments");
        g->e->set_tuning(*t);
    });
}

int mct_biome_tie_stats(mct_generator* g, int64_t out[4]) {
    return guard([&] {
        if (!g || !out) throw Error(MCT_EINVAL, "bad arguments");
        g->e->finish_all(nullptr);
        unsigned long long s[4];
        hip_check(hipMemcpy(s, g->e->d_stats, sizeof s, hipMemcpyDeviceToHost), "download stats");
        for (int k = 0; k < 4; ++k) out[k] = static_cast<int64_t>(s[k]);
    });
}

int mct_last_batch_done_ms(mct_generator* g, double* out, int cap, int* n) {
    return guard([&] {
        if (!g || !n || cap < 0 || (cap > 0 && !out)) throw Error(MCT_EINVAL, "bad arguments");
        const auto& d = g->e->done_ms;
        *n = static_cast<int>(d.size());
        for (int i = 0; i < cap && i < static_cast<int>(d.size()); ++i) out[i] = d[static_cast<size_t>(i)];
    });
}

void* mct_alloc_pinned(size_t bytes) {
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        g_err = "pinned host allocation failed";
        return nullptr;
    }
    return p;
}

void mct_free_pinned(void* p) { (void)hipHostFree(p); }

void* mct_alloc_device(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        g_err = "device allocation failed";
        return nullptr;
    }
    return p;
}

void mct_free_device(void* p) { (void)hipFree(p); }

int mct_device_to_host(void* dst, const void* src, size_t bytes) {
    return guard([&] { hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "device to host"); });
}

int mct_topdown(const uint8_t* terrain, int count, int on_device, uint8_t* records) {
    return guard([&] {
        if (count < 0 || (count > 0 && (!terrain || !records))) throw Error(MCT_EINVAL, "bad arguments");
        if (count == 0) return;
        constexpr int kGroup = 1 << 14;
        uint8_t *d_in = nullptr, *d_out = nullptr;
        try {
            if (!on_device) hip_check(hipMalloc(&d_in, static_cast<size_t>(std::min(count, kGroup)) * kPer), "malloc");
            hip_check(hipMalloc(&d_out, static_cast<size_t>(std::min(count, kGroup)) * 256 * 8), "malloc");
            for (int c0 = 0; c0 < count; c0 += kGroup) {
                const int nc = std::min(kGroup, count - c0);
                const uint8_t* src = terrain + static_cast<size_t>(c0) * kPer;
                if (!on_device) {
                    hip_check(hipMemcpy(d_in, src, static_cast<size_t>(nc) * kPer, hipMemcpyHostToDevice), "upload");
                    src = d_in;
                }
                k_topdown<<<blocks_for(static_cast<size_t>(nc) * 256, 256), 256>>>(src, nc, d_out);
                hip_check(hipGetLastError(), "topdown kernel");
                hip_check(hipMemcpy(records + static_cast<size_t>(c0) * 256 * 8, d_out, static_cast<size_t>(nc) * 256 * 8, hipMemcpyDeviceToHost), "download");
            }
        } catch (...) {
            (void)hipFree(d_in);
            (void)hipFree(d_out);
            throw;
        }
        (void)hipFree(d_in);
        hip_check(hipFree(d_out), "free");
    });
}

int mct_chunk_digests(const void* d_data, int64_t count, int64_t bytes_per_chunk, uint64_t* digests) {
    return guard([&] {
        if (count < 0 || (count > 0 && (!d_data || !digests)) || bytes_per_chunk <= 0 || bytes_per_chunk % 8 != 0 ||
            reinterpret_cast<uintptr_t>(d_data) % 8 != 0)
            throw Error(MCT_EINVAL, "bad arguments (bytes_per_chunk must be a positive multiple of 8, data 8-byte aligned)");
        if (count == 0) return;
        constexpr int64_t kGroup = int64_t{1} << 20;
        uint64_t* d = nullptr;
        hip_check(hipMalloc(&d, sizeof(uint64_t) * static_cast<size_t>(std::min(count, kGroup))), "malloc");
        try {
            for (int64_t c0 = 0; c0 < count; c0 += kGroup) {
                const int64_t nc = std::min(kGroup, count - c0);
                k_chunk_digest<<<static_cast<unsigned>(nc), 256>>>(static_cast<const uint64_t*>(d_data) + c0 * (bytes_per_chunk / 8), bytes_per_chunk / 8, d);
                hip_check(hipGetLastError(), "digest kernel");
                hip_check(hipMemcpy(digests + c0, d, sizeof(uint64_t) * static_cast<size_t>(nc), hipMemcpyDeviceToHost), "download");
            }
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        hip_check(hipFree(d), "free");
    });
}

int mct_chunk_digests_noalloc(const void* d_data, int64_t count, int64_t bytes_per_chunk, uint64_t* digests) {
    return guard([&] {
        if (count < 0 || (count > 0 && (!d_data || !digests)) || bytes_per_chunk <= 0 || bytes_per_chunk % 8 != 0 ||
            reinterpret_cast<uintptr_t>(d_data) % 8 != 0)
            throw Error(MCT_EINVAL, "bad arguments (bytes_per_chunk must be a positive multiple of 8, data 8-byte aligned)");
        if (count > MCT_DIGESTS_MAX)
            throw Error(MCT_ECAPACITY, "count " + std::to_string(count) + " beyond the fixed digest buffer (MCT_DIGESTS_MAX = " + std::to_string(MCT_DIGESTS_MAX) + ")");
        uint64_t* d = scratch_on_current_device();
        if (count == 0) return;
        k_chunk_digest<<<static_cast<unsigned>(count), 256>>>(static_cast<const uint64_t*>(d_data), bytes_per_chunk / 8, d);
        hip_check(hipGetLastError(), "digest kernel");
        hip_check(hipMemcpy(digests, d, sizeof(uint64_t) * static_cast<size_t>(count), hipMemcpyDeviceToHost), "download");
    });
}

int mct_profile(mct_generator* g, int enable) {
    return guard([&] {
        if (!g) throw Error(MCT_EINVAL, "null generator");
        g->e->finish_all(nullptr);
        g->e->profile = enable != 0;
        for (int k = 0; k < kKCount; ++k) {
            g->e->kernel_ms[k] = 0;
            g->e->kernel_launches[k] = 0;
        }
    });
}

int mct_kernel_times(mct_generator* g, double* ms, int64_t* launches, int cap, int* n) {
    return guard([&] {
        if (!g || !n) throw Error(MCT_EINVAL, "bad arguments");
        g->e->finish_all(nullptr);
        *n = kKCount;
        for (int k = 0; k < cap && k < kKCount; ++k) {
            if (ms) ms[k] = g->e->kernel_ms[k];
            if (launches) launches[k] = g->e->kernel_launches[k];
        }
    });
}

int mct_set_launch_mode(mct_generator* g, int graphs) {
    return guard([&] {
        if (!g) throw Error(MCT_EINVAL, "null generator");
        g->e->finish_all(nullptr);
        g->e->use_graphs = graphs != 0;
    });
}

const char* mct_kernel_name(int i) { return (i >= 0 && i < kKCount) ? kKernelNames[i] : ""; }

int mct_info(int32_t out[6]) {
    if (!out) return fail(MCT_EINVAL, "null out");
    out[0] = kPer;
    out[1] = kQPer;
    out[2] = kMinY;
    out[3] = kHeight;
    out[4] = mctgen::kPaletteSize;
    out[5] = mctgen::kNumBiomes;
    return MCT_OK;
}

}  // extern "C"
