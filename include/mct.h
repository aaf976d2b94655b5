/* C ABI of the MI355X TERRAIN-step generator (libmct.so).
 *
 * Drop-in boundary for the reference's GPU generator API, FullTerrainGenerator
 * (cuda/include/mccuda_cuda/terrain_gpu.h:31-63 in bonquifo/minecraft-terrain-cuda):
 * a seed and a chunk list go in; per chunk a u8 palette index per block and a u8 biome
 * index per quart come out. Layouts are the reference's (terrain_gpu.h:41-43):
 *   blocks  ((y - minY) * 16 + z) * 16 + x, 16 x 384 x 16 = 98,304 bytes per chunk
 *   biomes  (qy * 4 + qz) * 4 + qx, 4 x 96 x 4 = 1,536 bytes per chunk
 * chunks back to back in input order. Values index the terrain model's palette / biome list.
 *
 * Every entry point returns 0 on success and a non-zero status on failure; the message is
 * in mct_last_error() (thread-local). No C++ exception crosses this boundary. A generator
 * handle is not thread-safe: one host thread per handle (as the reference class).
 * The IR and terrain model are compiled into the library by the code generator; mct_create
 * checks that the given files are the ones it was generated from.
 */
#ifndef MCT_H
#define MCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MCT_OK = 0,
    MCT_EINVAL = 1,       /* bad argument (null handle, count < 0, unknown flag) */
    MCT_EMODEL = 2,       /* IR / terrain model files differ from the compiled-in model */
    MCT_EDEVICE = 3,      /* HIP runtime error (message carries the HIP error string) */
    MCT_ECAPACITY = 4,    /* request beyond a configured capacity */
    MCT_EIO = 5           /* cannot read a file */
};

typedef struct mct_generator mct_generator;

/* Per-stage times accumulated over a call (ms, HIP events on the batch streams), plus host
 * wall time; mirrors TerrainTiming (terrain_gpu.h:15-22). */
typedef struct mct_timing {
    double h2d_ms, biome_ms, noise_ms, aquifer_ms, fill_ms, surface_ms, d2h_ms, wall_ms;
} mct_timing;

/* FullTerrainGenerator(ir, tm, max_batch_chunks, pipeline_slots) (terrain_gpu.h:39).
 * ir_json / terrain_json: paths of the mccuda-ir/1 and mccuda-terrain/1 files.
 * max_batch_chunks: chunks per device batch (>= 1); pipeline_slots: batches in flight (1..4).
 * seed_capacity: per-seed noise tables kept resident on the device (>= 1; 0 = default 64). */
int mct_create(const char* ir_json, const char* terrain_json, int max_batch_chunks, int pipeline_slots, int seed_capacity, mct_generator** out);
int mct_destroy(mct_generator* g);

/* FullTerrainGenerator::set_seed (terrain_gpu.h:42): the seed of later mct_generate calls.
 * The seed's tables are built on the device (or found resident) at the next generate. */
int mct_set_seed(mct_generator* g, int64_t seed);

/* FullTerrainGenerator::generate (terrain_gpu.h:45): chunk_xz = count (x, z) pairs of chunk
 * coordinates. terrain (count * 98,304 B) is required; biomes (count * 1,536 B) and aquifer
 * (count * 98,304 B, the fill before material rules) may be NULL. Host memory; pinned
 * memory from mct_alloc_pinned is faster. Blocks until all results are in place. */
int mct_generate(mct_generator* g, const int32_t* chunk_xz, int count, uint8_t* terrain, uint8_t* biomes, uint8_t* aquifer, mct_timing* timing);

/* Batched form (no reference counterpart; the reference runs one seed per call): chunk i
 * uses seeds[i]. Any mix of seeds and uneven per-seed chunk lists; results do not depend on
 * the mix. Seeds beyond seed_capacity distinct values per call are processed in groups. */
int mct_generate_multi(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, uint8_t* terrain, uint8_t* biomes,
                       uint8_t* aquifer, mct_timing* timing);

/* Device-output form: terrain / biomes are device pointers (HBM, e.g. from mct_alloc_device);
 * nothing is copied to the host. Used to measure the generator with its outputs resident. */
int mct_generate_device(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, uint8_t* d_terrain, uint8_t* d_biomes,
                        mct_timing* timing);

/* Core configuration: TerrainGenerator::generate(chunks, count, density_out, fill_out, timing)
 * (cuda/include/mccuda_cuda/generator.h:30-45; CPU semantics cpu/src/volume_eval.cpp:676-690):
 * final density (f32 per block, the same layout as terrain) and the core fill (u8 per block:
 * 0 air, 1 stone, 2 water, 3 lava; no aquifers, no material rules, no biomes) of each chunk under
 * seeds[i]. density (count * 98,304 floats) or fill (count * 98,304 B) may be NULL, not both.
 * Host memory; blocks until all results are in place. */
int mct_generate_core(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, float* density, uint8_t* fill, mct_timing* timing);
/* The same with device-memory outputs (nothing copied to the host). */
int mct_generate_core_device(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, float* d_density, uint8_t* d_fill,
                             mct_timing* timing);

/* FullTerrainGenerator::biome_tie_stats (terrain_gpu.h:47): out[0] quarts searched, out[1]
 * lookups whose leaf the R-tree's last-result rule changed, out[2] of those, a changed biome.
 * out[3]: chunks whose surface pass was redone in exact sequential column order. */
int mct_biome_tie_stats(mct_generator* g, int64_t out[4]);

/* FullTerrainGenerator::last_batch_done_ms (terrain_gpu.h:51): per batch of the last call,
 * ms after the call started when its results were in place. Writes min(cap, batches) values
 * and stores the batch count in *n. */
int mct_last_batch_done_ms(mct_generator* g, double* out, int cap, int* n);

/* alloc_pinned / free_pinned (terrain_gpu.h:61-62), plus device memory for the device form. */
void* mct_alloc_pinned(size_t bytes);
void mct_free_pinned(void* p);
void* mct_alloc_device(size_t bytes);
void mct_free_device(void* p);
int mct_device_to_host(void* dst, const void* src, size_t bytes);

/* Top-down view records of count chunks of terrain (device memory when on_device != 0, else
 * host memory): per chunk 256 records in (z * 16 + x) order, 8 bytes each: top state (highest
 * non-air block), floor state (first non-fluid block under it), min(255, top_y - floor_y), 0,
 * int16 top y, int16 floor y (little-endian); the column record of the reference bench's MCTD
 * file (bench/bench_main.cpp:132-166). records: count * 2,048 bytes, host memory. */
int mct_topdown(const uint8_t* terrain, int count, int on_device, uint8_t* records);

/* Per-chunk 64-bit digests of count chunks of bytes_per_chunk bytes each at d_data (device
 * memory, e.g. the terrain or biome output of mct_generate_device), computed on the device; no
 * reference counterpart. digest = mix(sum over 8-byte little-endian words w_i, i = 0.., of
 * mix(w_i + (i + 1) * 0x9E3779B97F4A7C15) mod 2^64, xor bytes_per_chunk), mix = SplitMix64's
 * finaliser. bytes_per_chunk: a positive multiple of 8. digests: count values, host memory. */
int mct_chunk_digests(const void* d_data, int64_t count, int64_t bytes_per_chunk, uint64_t* digests);
/* mct_chunk_digests without any device allocation, for callers that cannot allocate (tight
 * serving loops, harnesses that forbid allocation). No counterpart in the reference. The digests
 * pass through a fixed device buffer of the library (MCT_DIGESTS_MAX values per GPU, held until
 * the process ends) instead of a per-call allocation, so count is at most MCT_DIGESTS_MAX
 * (MCT_ECAPACITY beyond). Calls on one GPU share that buffer: calls that overlap on one GPU
 * (from several host threads) race. Same digests as mct_chunk_digests. */
#define MCT_DIGESTS_MAX 65536
int mct_chunk_digests_noalloc(const void* d_data, int64_t count, int64_t bytes_per_chunk, uint64_t* digests);

/* ---- Launch shapes and the autotuner (no counterpart in the reference API; the reference picks
 * its tiles with a separate tuning program) ----
 * workgroup[k]: workgroup size of kernel group k (mct_kernel_name(k)); 64, 128 or 256 where the
 * group's kernels take any size, 0 where the kernel's layout fixes it (and 0 in mct_set_tuning
 * also means the default, 256). run_len: quarts per R-tree search thread (1..96, default 8).
 * flags: MCT_TUNE_REPLAY_ALL makes the exact surface pass walk every column of each chunk it
 * redoes again, in place of only the columns whose steep test the final heights change (a check
 * of the selective replay); other bits must be 0. ms[k]: from mct_autotune, group k's time on
 * the sample with the chosen shape. Results never depend on the shapes or the flags. */
#define MCT_TUNE_REPLAY_ALL 1
typedef struct mct_tuning {
    int32_t workgroup[24];
    int32_t run_len;
    int32_t flags;
    double ms[24];
} mct_tuning;
/* Times the shapes on a sample (its first max_batch_chunks chunks; chunk i under seeds[i]; outputs
 * are discarded), keeps the fastest for each knob and writes them to out (may be NULL). */
int mct_autotune(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, mct_tuning* out);
int mct_get_tuning(mct_generator* g, mct_tuning* out);
/* Applies a stored tuning (e.g. from an earlier mct_autotune on the same device and model). */
int mct_set_tuning(mct_generator* g, const mct_tuning* t);

/* Per-kernel timing (HIP events around each kernel group on its stream; adds a little
 * overhead, off by default). enable != 0 turns it on and clears the totals. */
int mct_profile(mct_generator* g, int enable);
/* Accumulated ms and launch counts per kernel group since mct_profile; *n = group count. */
int mct_kernel_times(mct_generator* g, double* ms, int64_t* launches, int cap, int* n);
/* How untimed batches are launched: graphs != 0 (the default) replays one captured HIP graph
 * per batch shape, graphs == 0 launches the same kernels one by one. Results are identical.
 * No counterpart in the reference (its batches are stream launches, terrain.cu:1292-1336). */
int mct_set_launch_mode(mct_generator* g, int graphs);
const char* mct_kernel_name(int i);

/* Biome-only queries (no counterpart in the reference API): for block i at xyz[3i..3i+2] under
 * seeds[i], the biome index the TERRAIN step stores for that block (BiomeManager zoom, then the
 * quart's biome as left by the chunk's chain of climate R-tree searches), without generating
 * chunks. biomes: count bytes. |x|, |z| <= 2^30 (MCT_EINVAL beyond); any y (quart y clamps to
 * the world height). MCT_ECAPACITY if the model's climate bounds or a query's climate targets
 * exceed the float tree's exact range (2^23). */
int mct_biome_query(mct_generator* g, const int64_t* seeds, const int32_t* xyz, int count, uint8_t* biomes);

/* zoom_selftest (terrain_gpu.h:57-59): the column-cached zoom that the surface pass uses against
 * the plain FP64 BiomeManager zoom, on about `count` blocks (whole columns at pseudo-random
 * x, z in +-29,999,984, every y). out: blocks checked, mismatches, near ties (the reference's
 * "fallbacks": picks with a second candidate within 8e-6 of the minimal fiddled distance, which
 * the reference's FP32 filter hands to FP64; here every distance is FP64). */
int mct_zoom_selftest(int64_t world_seed, int64_t count, int64_t out[3]);

/* ---- Serving (no counterpart in the reference, which serves one synchronous call at a time) ----
 * A server owns a generator with two pipeline slots, the per-seed tables of up to
 * seed_capacity seeds, and a result arena in HBM sized by hbm_budget_bytes (the budget minus
 * the generator's own buffers). Region requests are admitted, in arrival order, while the
 * arena has room for their outputs. A scheduler thread batches the admitted requests' chunks
 * in arrival order (one batch mixes requests and seeds; a request arriving while others run
 * joins the next batch), writes results into the arena, and copies each finished request to
 * its host buffers on a separate stream while later batches compute. Results are the same
 * bytes as mct_generate gives for the request alone. Submit / poll / drain may be called from
 * any thread. */
typedef struct mct_server mct_server;

/* max_batch_chunks: chunks per device batch (1..65536); seed_capacity as in mct_create. */
int mct_server_create(const char* ir_json, const char* terrain_json, int max_batch_chunks, int seed_capacity, size_t hbm_budget_bytes,
                      mct_server** out);
/* Queues a request: count chunks of seed; terrain (count * 98,304 B) and biomes (count * 1,536
 * B, may be NULL) are host buffers (pinned is faster) written before the request completes.
 * MCT_ECAPACITY if the request alone exceeds the arena. */
int mct_server_submit(mct_server* s, int64_t seed, const int32_t* chunk_xz, int count, uint8_t* terrain, uint8_t* biomes, int64_t* request_id);
/* Up to cap requests completed since the last poll: their ids and their latency (ms from
 * submit to the end of the copy to the host); *n = how many. */
int mct_server_poll(mct_server* s, int64_t* ids, double* latency_ms, int cap, int* n);
/* Blocks until every submitted request has completed. */
int mct_server_drain(mct_server* s);
/* submitted, completed, batches, chunks, arena chunk slots, most slots held at once, batches
 * that mixed requests, seed capacity. */
int mct_server_stats(mct_server* s, int64_t out[8]);
int mct_server_destroy(mct_server* s);

/* Build facts: blocks per chunk, quarts per chunk, min y, height, palette size, biome count. */
int mct_info(int32_t out[6]);

const char* mct_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MCT_H */
