// Host engine of the TERRAIN step on one GPU.
//
// A batch (at most max_batch chunks, at most seed_capacity distinct seeds) is one fixed device
// program: the H2D copy of a pinned staging block (the batch descriptor, the seed-table jobs,
// the chunk list), the seed-table builds, the biome-chunk union, biomes, noise corners,
// surface grid and aquifer cells, fill and material rules. Every per-batch value lives in the
// staging block or in device counts, so the program's launches have fixed arguments: it is
// captured once per shape as a graph (hipStreamBeginCapture .. hipGraphInstantiate) and
// replayed with one hipGraphLaunch per batch. The shape is what the launch sequence depends
// on: the radix passes of the three dedupes (the key widths), which optional outputs are
// written, the core or full configuration and the launch-shape tuning. Timed batches (kernel
// or stage times asked for) run the same program launch by launch with events around groups.
//
// Lanes: each lane is a stream with its own device buffers, staging block and graphs; batches
// go to the lanes in turn. Host outputs alternate between two result buffers per lane and are
// copied on one copy stream shared by the lanes, so a lane's next batch computes while its
// previous results are copied; a call whose results go to host memory is cut into smaller
// batches, so that the copies start early and run back to back.
//
// Seed tables: seed_capacity resident tables in HBM, least recently used replaced. A batch
// lists the tables it needs built as jobs, and its program builds them first. A table is only
// replaced once the batches that read it have finished, and a batch reading a table another
// lane's batch is still building waits for that batch.
#pragma once

#include <map>
#include <memory>
#include <unordered_map>
#include <vector>

#include "dedupe.h"

namespace mct {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(MCT_EDEVICE, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

inline unsigned blocks_for(size_t threads, int block) { return static_cast<unsigned>((threads + block - 1) / block); }

// Kernel groups: timed with HIP events in timed batches, and the units of the launch-shape
// tuning.
enum KernelId {
    kKBiomeColumns, kKBiomeQuarts, kKBiomeTies, kKBiomeChain, kKDedupeCols, kKColumns, kKCorners, kKDedupeGrid, kKFts, kKGridScatter,
    kKAqCells, kKFill, kKHeights, kKSurface1, kKSurface2, kKSurfaceExact, kKGather, kKSurfaceCols, kKOre, kKSeedSetup, kKBiomeUnion, kKCount
};
inline const char* const kKernelNames[kKCount] = {"biome_columns", "biome_quarts", "biome_ties", "biome_chain", "dedupe_columns", "columns", "corners",
                                                  "dedupe_grid", "fts", "grid_scatter+aquifer_skip", "aquifer_cells", "fill", "heights", "surface1",
                                                  "surface2", "surface_check+exact", "gather_biomes", "surface_columns", "ore_veins", "seed_setup",
                                                  "biome_union"};

// Stages of mct_timing, bounded by events in timed batches.
enum Stage { kStH2D, kStBiome, kStNoise, kStAquifer, kStFill, kStSurface, kStCount };

// The seed-independent model on the device: R-tree (int64 and float forms), leaf tables,
// the blended-noise layer starts, the constant-memory model record.
struct DeviceModel {
    TreeNodeDev* tree = nullptr;
    TreeNodeF* treef = nullptr;
    uint8_t* leaf_biome = nullptr;
    int32_t* leaf_node = nullptr;
    int32_t* leaf_rank = nullptr;
    int* blended_first = nullptr;
    int vary_groups = 0;        // column-search groups of the one varying climate coordinate (0: none)
    int nnormal = 0;            // NormalNoise layers (the first layers of the table)

    DeviceModel() {
        std::vector<TreeNodeDev> t(mctgen::kTreeNodes);
        for (int i = 0; i < mctgen::kTreeNodes; ++i) {
            const mctgen::TreeNode& s = mctgen::kTree[i];
            for (int k = 0; k < 7; ++k) { t[i].lo[k] = s.lo[k]; t[i].hi[k] = s.hi[k]; }
            t[i].first = s.first;
            t[i].count = s.count;
            t[i].leaf = s.leaf;
            t[i].pad = 0;
        }
        upload(&tree, t.data(), t.size());
        const std::vector<TreeNodeF> tf = float_tree(t);
        if (!tf.empty() && tree_levels(t) <= kSearchDepth) upload(&treef, tf.data(), tf.size());
        upload(&leaf_biome, mctgen::kLeafBiome, mctgen::kTreeLeaves);
        upload(&leaf_node, mctgen::kLeafNode, mctgen::kTreeLeaves);
        const std::vector<int32_t> rank = leaf_dfs_rank(t, mctgen::kTreeLeaves);
        upload(&leaf_rank, rank.data(), rank.size());
        ModelDev M{};
        M.tree = tree;
        M.treef = treef;
        M.leaf_biome = leaf_biome;
        M.leaf_node = leaf_node;
        M.leaf_rank = leaf_rank;
        if (treef) column_groups(tf, M);
        vary_groups = (treef && kOneVarying) ? M.col_groups[kVaryCoord] : 0;
        std::memcpy(M.block_flags, mctgen::kBlockFlags, 64);
        std::memcpy(M.biome_temp, mctgen::kBiomeTemp, sizeof M.biome_temp);
        std::memcpy(M.biome_frozen, mctgen::kBiomeFrozen, sizeof M.biome_frozen);
        std::memcpy(M.simplex, mctgen::kSimplexPerm, sizeof M.simplex);
        for (int k = 0; k < 7; ++k) M.terracotta[k] = mctgen::kTerracotta[k];
        for (int k = 0; k < mctgen::kNumOreRules; ++k) {
            for (int j = 0; j < 3; ++j) M.ore_state[k][j] = mctgen::kOreState[k][j];
            M.ore_raw[k] = mctgen::kOreRawChance[k];
        }
        hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_model), &M, sizeof M), "upload model");
        std::vector<int> bf(std::max(mctgen::kNumBlended, 1), -1);
        for (int l = 0; l < mctgen::kNumLayers; ++l) {
            const mctgen::LayerRecipe& r = mctgen::kLayerRecipe[l];
            if (r.kind == 0) nnormal = l + 1;
            else if (bf[r.a] < 0) bf[r.a] = l;
        }
        upload(&blended_first, bf.data(), bf.size());
    }
    ~DeviceModel() {
        void* p[] = {tree, treef, leaf_biome, leaf_node, leaf_rank, blended_first};
        for (void* q : p) (void)hipFree(q);
    }
    DeviceModel(const DeviceModel&) = delete;
    DeviceModel& operator=(const DeviceModel&) = delete;

    template <class T>
    static void upload(T** dst, const T* src, size_t count) {
        hip_check(hipMalloc(reinterpret_cast<void**>(dst), sizeof(T) * std::max<size_t>(count, 1)), "malloc model");
        hip_check(hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice), "upload model");
    }
};

// Resident seed tables: which seed each table holds, the batch that last read it and the
// batch (and lane) that built it.
struct SeedCache {
    int cap;
    SeedDev* d = nullptr;
    std::vector<int64_t> seed;
    std::vector<int64_t> last_use, built_by;   // batch sequence numbers (-1: none)
    std::vector<int> built_lane;
    std::unordered_map<int64_t, int> where;

    explicit SeedCache(int capacity) : cap(capacity), seed(capacity, 0), last_use(capacity, -1), built_by(capacity, -1), built_lane(capacity, -1) {
        hip_check(hipMalloc(&d, sizeof(SeedDev) * static_cast<size_t>(cap)), "malloc seed tables");
    }
    ~SeedCache() { (void)hipFree(d); }
    SeedCache(const SeedCache&) = delete;
    SeedCache& operator=(const SeedCache&) = delete;
};

// What one batch program looks like (the graph cache key): the chunk-count bucket the
// one-shot kernels' grids are sized for, the radix passes of the three dedupes, the outputs.
// A dedupe whose pass count is kDensePass takes the dense form (its key space fits the lane's
// dense array) instead of the radix sort.
constexpr int kDensePass = -1;

struct Shape {
    bool core = false;
    int grid_n = 0;
    int pass_bio = 0, pass_col = 0, pass_grid = 0;
    bool biomes = false, aquifer = false, density = false, fill = true;
    bool operator<(const Shape& o) const {
        const int a[] = {core, grid_n, pass_bio, pass_col, pass_grid, biomes, aquifer, density, fill};
        const int b[] = {o.core, o.grid_n, o.pass_bio, o.pass_col, o.pass_grid, o.biomes, o.aquifer, o.density, o.fill};
        return std::lexicographical_compare(a, a + 9, b, b + 9);
    }
};

// Where a batch's results go: device pointers the kernels write, or host pointers the lane's
// copies fill from its buffers.
struct BatchOut {
    uint8_t *terrain = nullptr, *biomes = nullptr, *aquifer = nullptr;
    float* density = nullptr;
    bool device = false;
};

// Device memory of one lane, sized for `cap` chunks (biome chunks: at most 9 per chunk).
constexpr size_t kDescBytes = 1024;   // the batch descriptor's slot at the start of the staging block

struct LaneMemory {
    int cap;
    size_t bytes = 0;
    std::vector<void*> owned;
    // staging block: descriptor, seed jobs, chunks (pinned host and device copies)
    size_t block_bytes = 0, jobs_off = 0, chunks_off = 0;
    uint8_t *h_block = nullptr, *d_block = nullptr;
    // biome chunks
    uint64_t *bkeys = nullptr, *ubk = nullptr;
    int* nb = nullptr;
    ChunkRef* bch = nullptr;
    float* bcol = nullptr;
    int64_t *targets = nullptr, *mind = nullptr, *tvary = nullptr, *tcol = nullptr;
    Targets tgt() const { return Targets{targets, tcol, tvary}; }
    int32_t* lstar = nullptr;
    uint8_t *tie = nullptr, *bb = nullptr;
    uint64_t* bmask = nullptr;
    GroupRes* gres = nullptr;
    // corner columns, surface grid, aquifer cells
    uint64_t *ckeys = nullptr, *ucols = nullptr, *gkeys = nullptr, *ug = nullptr;
    int *colmap = nullptr, *gmap = nullptr, *skip = nullptr, *counts = nullptr, *flags = nullptr;
    float *colv = nullptr, *corners = nullptr, *gv = nullptr, *prelim = nullptr, *colrec = nullptr;
    int32_t *grid = nullptr, *cloc = nullptr, *cfl = nullptr;
    // outputs and surface state
    uint8_t *blocks = nullptr, *bout = nullptr, *aq = nullptr, *ore = nullptr;
    uint8_t *blocks2 = nullptr, *bout2 = nullptr;   // second result buffers (host outputs)
    float* density = nullptr;
    int *h0 = nullptr, *h1 = nullptr, *h2 = nullptr, *hf = nullptr;
    uint32_t* clsw = nullptr;
    // dedupe scratch (capacity: the surface grid's keys, the largest of the three)
    int dcap = 0;
    uint64_t *dka = nullptr, *dkb = nullptr;
    int *dva = nullptr, *dvb = nullptr, *dflag = nullptr, *dhist = nullptr, *dhsum = nullptr, *dtsum = nullptr, *dtotal = nullptr;
    // the dense form's flag / rank array (dense_cap ints) and its tile sums
    size_t dense_cap = 0;
    int *dense_f = nullptr, *dense_t = nullptr;

    LaneMemory(int c, int seed_cap) : cap(c) {
        jobs_off = kDescBytes;
        chunks_off = jobs_off + ((sizeof(SeedJob) * static_cast<size_t>(seed_cap) + 255) & ~size_t{255});
        block_bytes = chunks_off + sizeof(ChunkRef) * static_cast<size_t>(cap);
        static_assert(sizeof(BatchDesc) <= kDescBytes, "the batch descriptor fits its slot of the staging block");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_block), block_bytes, hipHostMallocDefault), "pinned staging");
        std::memset(h_block, 0, block_bytes);
        M(&d_block, block_bytes);
        const size_t nbk = 9 * static_cast<size_t>(cap), ncol = std::max(mctgen::kNumColumns, 1), ncc = std::max(mctgen::kNumClimateColumns, 1);
        M(&bkeys, nbk);
        M(&ubk, nbk);
        M(&nb, nbk);
        M(&bch, nbk);
        M(&bcol, nbk * 16 * ncc);
        M(&targets, kOneVarying ? 0 : nbk * kQPer * 7);
        M(&tcol, kOneVarying ? nbk * 16 * 7 : 0);
        M(&tvary, nbk * kQPer);
        M(&mind, nbk * kQPer);
        M(&lstar, nbk * kQPer);
        M(&tie, nbk * kQPer);
        M(&bb, nbk * kQPer);
        M(&bmask, nbk * 16 * 2);
        M(&gres, nbk * 16 * kColGroups);
        const size_t ck = static_cast<size_t>(cap) * kCols, gk = static_cast<size_t>(cap) * kGridKeys;
        M(&ckeys, ck);
        M(&ucols, ck);
        M(&colmap, ck);
        M(&colv, ck * ncol);
        M(&corners, ck * kCh * kRows);
        M(&gkeys, gk);
        M(&ug, gk);
        M(&gmap, gk);
        M(&gv, gk);
        M(&grid, static_cast<size_t>(cap) * kGridCells);
        M(&prelim, static_cast<size_t>(cap) * 4);
        M(&skip, static_cast<size_t>(cap));
        M(&cloc, static_cast<size_t>(cap) * kAqCells * 4);
        M(&cfl, static_cast<size_t>(cap) * kAqCells * 2);
        M(&blocks, static_cast<size_t>(cap) * kPer);
        M(&bout, static_cast<size_t>(cap) * kQPer);
        M(&h0, static_cast<size_t>(cap) * 256);
        M(&h1, static_cast<size_t>(cap) * 256);
        M(&h2, static_cast<size_t>(cap) * 256);
        M(&hf, static_cast<size_t>(cap) * 256);
        M(&colrec, static_cast<size_t>(cap) * 256 * kColF);
        M(&ore, static_cast<size_t>(cap) * 256 * kOreRowsPad);
        M(&clsw, static_cast<size_t>(cap) * (kHeight / 16) * 256);
        M(&flags, static_cast<size_t>(cap));
        M(&counts, kCntSlots);
        dcap = static_cast<int>(gk);
        const size_t tiles = (static_cast<size_t>(dcap) + kRsTile - 1) / kRsTile;
        M(&dka, static_cast<size_t>(dcap));
        M(&dkb, static_cast<size_t>(dcap));
        M(&dva, static_cast<size_t>(dcap));
        M(&dvb, static_cast<size_t>(dcap));
        M(&dflag, static_cast<size_t>(dcap));
        M(&dhist, 256 * tiles);
        M(&dhsum, (256 * tiles + kScanTile - 1) / kScanTile + 1);
        M(&dtsum, tiles + 1);
        M(&dtotal, 1);
        // four ints per grid key of capacity (at most 2^26): the spaces of contiguous batches
        // (a 64 x 64 region's grid keys span about 150K) with room to spare
        dense_cap = std::min<size_t>(4 * gk, size_t{1} << 26);
        M(&dense_f, dense_cap);
        M(&dense_t, dense_cap / kDenseTile + 1);
    }
    ~LaneMemory() {
        for (void* p : owned) (void)hipFree(p);
        (void)hipHostFree(h_block);
    }
    LaneMemory(const LaneMemory&) = delete;
    LaneMemory& operator=(const LaneMemory&) = delete;

    template <class T>
    void M(T** p, size_t count) {
        const size_t b = sizeof(T) * std::max<size_t>(count, 1);
        hip_check(hipMalloc(reinterpret_cast<void**>(p), b), "malloc lane buffers");
        owned.push_back(*p);
        bytes += b;
    }
    template <class T>
    void ensure(T** p, size_t count) {
        if (!*p) M(p, count);
    }

    BatchDesc* h_desc() const { return reinterpret_cast<BatchDesc*>(h_block); }
    SeedJob* h_jobs() const { return reinterpret_cast<SeedJob*>(h_block + jobs_off); }
    ChunkRef* h_chunks() const { return reinterpret_cast<ChunkRef*>(h_block + chunks_off); }
    const BatchDesc* d_desc() const { return reinterpret_cast<const BatchDesc*>(d_block); }
    const SeedJob* d_jobs() const { return reinterpret_cast<const SeedJob*>(d_block + jobs_off); }
    const ChunkRef* d_chunks() const { return reinterpret_cast<const ChunkRef*>(d_block + chunks_off); }
    const int* d_count_of(size_t field) const { return reinterpret_cast<const int*>(d_block + field); }
    const KeyCodec* d_codec(size_t field) const { return reinterpret_cast<const KeyCodec*>(d_block + field); }
};

struct Lane {
    std::unique_ptr<LaneMemory> mem;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr, copied = nullptr;   // the latest batch's program / copies done
    // Host outputs alternate between two result buffers per lane, so the lane's next batch
    // computes while the previous one's results are still being copied; buf_free[k]: the copy
    // that last read buffer k is done.
    int obuf = 0;
    hipEvent_t buf_free[2] = {};
    hipEvent_t stage_ev[kStCount + 1] = {}, kernel_ev[kKCount][2] = {}, copy_ev[2] = {};
    std::map<Shape, hipGraphExec_t> graphs;
    int64_t seq = -1;             // the lane's latest batch
    int batch_index = -1;         // its index in the current call (completion times)
    bool pending = false;         // launched, completion not yet collected
    bool timed = false, copies = false;

    Lane(int cap, int seed_cap) : mem(std::make_unique<LaneMemory>(cap, seed_cap)) {
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
        hip_check(hipEventCreateWithFlags(&done, hipEventDefault), "event");
        hip_check(hipEventCreateWithFlags(&copied, hipEventDefault), "event");
        for (auto& e : buf_free) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
        for (auto& e : stage_ev) hip_check(hipEventCreate(&e), "event");
        for (auto& e : copy_ev) hip_check(hipEventCreate(&e), "event");
        hip_check(hipEventRecord(done, stream), "event");
        hip_check(hipEventRecord(copied, stream), "event");
        for (auto& e : buf_free) hip_check(hipEventRecord(e, stream), "event");
    }
    ~Lane() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        for (auto& e : stage_ev) (void)hipEventDestroy(e);
        for (auto& e : copy_ev) (void)hipEventDestroy(e);
        for (auto& k : kernel_ev)
            for (auto& e : k)
                if (e) (void)hipEventDestroy(e);
        (void)hipEventDestroy(done);
        (void)hipEventDestroy(copied);
        for (auto& e : buf_free) (void)hipEventDestroy(e);
        mem.reset();
        if (stream) (void)hipStreamDestroy(stream);
    }
    Lane(const Lane&) = delete;
    Lane& operator=(const Lane&) = delete;
    void drop_graphs() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }
};

class Engine {
public:
    const int max_batch, seed_cap;
    DeviceModel model;
    SeedCache seeds;
    std::vector<std::unique_ptr<Lane>> lanes;
    unsigned long long* d_stats = nullptr;   // [0] quarts, [1] chain overrides, [2] biome changes, [3] exact redo
    // launch shapes: workgroup size per tunable kernel group, quarts per scalar R-tree search
    int wg[kKCount] = {};
    int run_len = kBiomeRunDefault;
    bool replay_all = false;   // exact surface pass: every column of a redone chunk
    bool use_graphs = true;
    bool profile = false;
    double kernel_ms[kKCount] = {};
    int64_t kernel_launches[kKCount] = {};
    std::vector<double> done_ms;
    std::vector<hipEvent_t> batch_ev;   // completion of batch i of the current call (results in place)

    // Copies of host outputs, for every lane (created with the first one): one stream beside
    // the lanes' keeps the process within the device's 4 hardware queues per process (streams
    // beyond that share a queue, and a stream's wait for a copy then stalls another's kernels).
    hipStream_t copy_stream = nullptr;

    Engine(int mb, int nlanes, int sc) : max_batch(mb), seed_cap(sc), model(), seeds(sc) {
        hip_check(hipGetDevice(&device), "get device");
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "device attribute");
        hip_check(hipMalloc(&d_stats, sizeof(unsigned long long) * 4), "malloc");
        hip_check(hipMemset(d_stats, 0, sizeof(unsigned long long) * 4), "memset");
        hip_check(hipEventCreate(&t0), "event");
        for (int i = 0; i < nlanes; ++i) lanes.push_back(std::make_unique<Lane>(mb, sc));
        hip_check(hipDeviceSynchronize(), "init");
    }
    ~Engine() {
        (void)hipDeviceSynchronize();
        lanes.clear();
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        for (hipEvent_t e : batch_ev) (void)hipEventDestroy(e);
        void* q[] = {d_qxyz, d_qslot, d_qbad, d_qout};
        for (void* x : q) (void)hipFree(x);
        (void)hipFree(d_stats);
        if (t0) (void)hipEventDestroy(t0);
    }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    int W(int k) const { return wg[k] > 0 ? wg[k] : 256; }

    size_t device_bytes() const {
        size_t b = sizeof(SeedDev) * static_cast<size_t>(seed_cap) + sizeof(TreeNodeDev) * mctgen::kTreeNodes + sizeof(TreeNodeF) * mctgen::kTreeNodes;
        for (const auto& l : lanes) b += l->mem->bytes;
        return b;
    }

    // ---- calls ------------------------------------------------------------------------------
    // The chunk list split into batches of at most max_batch chunks and seed_cap distinct
    // seeds, run on the lanes in turn; returns when every result is in place.
    void generate(const int64_t* sd, const int32_t* xz, int count, BatchOut out, mct_timing* tm, bool core = false) {
        const auto w0 = std::chrono::steady_clock::now();
        finish_all(nullptr);
        done_ms.clear();
        hip_check(hipEventRecord(t0, lanes[0]->stream), "event");
        // Results copied to host memory: the host link (about 57 GB/s, 580K chunks/s of terrain)
        // is slower than the kernels, so the call is cut into batches of about an eighth of it
        // (at least 1024 chunks), the first one half that, so that the copies start early and
        // then run back to back while the lanes compute the next batches.
        const bool to_host = !out.device && (out.terrain || out.biomes || out.aquifer || out.density);
        const int step = (to_host && lanes.size() > 1) ? std::min(max_batch, std::max(1024, (count + 7) / 8)) : max_batch;
        int index = 0;
        for (int off = 0; off < count; ++index) {
            int n = std::min(index == 0 && step < max_batch ? step / 2 : step, count - off);
            {
                std::unordered_map<int64_t, int> seen;
                for (int i = 0; i < n; ++i)
                    if (seen.emplace(sd[off + i], 1).second && static_cast<int>(seen.size()) > seed_cap) { n = i; break; }
            }
            BatchOut o = out;
            const size_t o8 = static_cast<size_t>(off);
            if (o.terrain) o.terrain += o8 * kPer;
            if (o.biomes) o.biomes += o8 * kQPer;
            if (o.aquifer) o.aquifer += o8 * kPer;
            if (o.density) o.density += o8 * kPer;
            Lane& L = *lanes[static_cast<size_t>(next_lane)];
            next_lane = (next_lane + 1) % static_cast<int>(lanes.size());
            finish(L, tm);
            done_ms.push_back(0.0);
            L.batch_index = index;
            launch(L, sd + off, xz + 2 * o8, n, o, core, tm != nullptr || profile);
            off += n;
        }
        finish_all(tm);
        if (tm) tm->wall_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    }

    // One batch on lane L, asynchronously (the serving loop's entry): n chunks, device outputs.
    void launch_device(int lane, const int64_t* sd, const int32_t* xz, int n, uint8_t* d_terrain, uint8_t* d_biomes) {
        Lane& L = *lanes[static_cast<size_t>(lane)];
        finish(L, nullptr);
        BatchOut o;
        o.terrain = d_terrain;
        o.biomes = d_biomes;
        o.device = true;
        L.batch_index = -1;
        launch(L, sd, xz, n, o, false, false);
    }
    bool lane_pending(int lane) const { return lanes[static_cast<size_t>(lane)]->pending; }
    bool lane_ready(int lane) const {
        const Lane& L = *lanes[static_cast<size_t>(lane)];
        return !L.pending || hipEventQuery(L.done) == hipSuccess;
    }
    void lane_finish(int lane) { finish(*lanes[static_cast<size_t>(lane)], nullptr); }

    // Every lane's batch and every copy done; the call's per-batch completion times.
    void finish_all(mct_timing* tm) {
        for (auto& l : lanes) finish(*l, tm);
        for (auto& l : lanes) hip_check(hipEventSynchronize(l->copied), "copies");
        for (size_t i = 0; i < done_ms.size() && i < batch_ev.size(); ++i) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, t0, batch_ev[i]) == hipSuccess) done_ms[i] = ms;
        }
    }

    // ---- launch shapes -----------------------------------------------------------------------
    static bool tunable(int k) {
        switch (k) {
            case kKBiomeColumns: case kKBiomeQuarts: case kKBiomeTies: case kKBiomeChain: case kKColumns: case kKFts: case kKAqCells:
            case kKFill: case kKHeights: case kKSurfaceCols: case kKOre: return true;
            default: return false;
        }
    }
    void set_shapes(const int* w, int rl) {
        finish_all(nullptr);
        for (int k = 0; k < kKCount; ++k) wg[k] = w[k];
        run_len = rl;
        for (auto& l : lanes) l->drop_graphs();
    }

    void get_tuning(mct_tuning* out) const {
        *out = mct_tuning{};
        for (int k = 0; k < kKCount; ++k) out->workgroup[k] = tunable(k) ? W(k) : 0;
        out->run_len = run_len;
        out->flags = replay_all ? MCT_TUNE_REPLAY_ALL : 0;
    }

    void set_tuning(const mct_tuning& t) {
        for (int k = 0; k < kKCount; ++k) {
            const int v = t.workgroup[k];
            if (tunable(k) ? !(v == 0 || v == 64 || v == 128 || v == 256) : v != 0)
                throw Error(MCT_EINVAL, std::string("workgroup size of ") + kKernelNames[k] + " must be " + (tunable(k) ? "0, 64, 128 or 256" : "0"));
        }
        if (t.run_len < 1 || t.run_len > kQH) throw Error(MCT_EINVAL, "run_len must be 1.." + std::to_string(kQH));
        if (t.flags & ~MCT_TUNE_REPLAY_ALL) throw Error(MCT_EINVAL, "unknown tuning flags");
        finish_all(nullptr);
        replay_all = (t.flags & MCT_TUNE_REPLAY_ALL) != 0;
        set_shapes(t.workgroup, t.run_len);
    }

    // Times each launch shape on a sample (its first max_batch chunks, outputs in HBM): per
    // tunable kernel group the workgroup sizes 64 / 128 / 256, then the quarts per scalar
    // R-tree search thread; keeps the fastest of each (coordinate descent, one knob at a time,
    // the group's own HIP-event time, best of two timed passes after a warm pass).
    void autotune(const int64_t* sd, const int32_t* xz, int count, mct_tuning* out) {
        count = std::min(count, max_batch);
        finish_all(nullptr);
        static_assert(kKCount <= 24, "mct_tuning holds 24 kernel groups");
        uint8_t* d_t = nullptr;
        hip_check(hipMalloc(&d_t, static_cast<size_t>(count) * kPer), "malloc");
        const bool was = profile;
        profile = true;
        double best_ms[kKCount] = {};
        BatchOut o;
        o.terrain = d_t;
        o.device = true;
        auto group_ms = [&](int k) {
            double best = 1e300;
            generate(sd, xz, count, o, nullptr);
            for (int rep = 0; rep < 2; ++rep) {
                for (int g = 0; g < kKCount; ++g) kernel_ms[g] = 0.0, kernel_launches[g] = 0;
                generate(sd, xz, count, o, nullptr);
                best = std::min(best, kernel_ms[k]);
            }
            return best;
        };
        try {
            const int sizes[] = {64, 128, 256};
            for (int k = 0; k < kKCount; ++k) {
                if (!tunable(k)) continue;
                int pick = W(k);
                double pick_ms = 1e300;
                for (int v : sizes) {
                    wg[k] = v;
                    const double ms = group_ms(k);
                    if (ms < pick_ms) { pick_ms = ms; pick = v; }
                }
                wg[k] = pick;
                best_ms[k] = pick_ms;
            }
            if (!model.treef) {
                const int runs[] = {4, 6, 8, 12, 16};
                int pick = run_len;
                double pick_ms = 1e300;
                for (int v : runs) {
                    run_len = v;
                    const double ms = group_ms(kKBiomeQuarts);
                    if (ms < pick_ms) { pick_ms = ms; pick = v; }
                }
                run_len = pick;
                best_ms[kKBiomeQuarts] = pick_ms;
            }
        } catch (...) {
            profile = was;
            (void)hipFree(d_t);
            throw;
        }
        profile = was;
        for (int g = 0; g < kKCount; ++g) kernel_ms[g] = 0.0, kernel_launches[g] = 0;
        hip_check(hipFree(d_t), "free");
        for (auto& l : lanes) l->drop_graphs();
        if (out) {
            get_tuning(out);
            for (int g = 0; g < kKCount; ++g) out->ms[g] = best_ms[g];
        }
    }

private:
    int device = 0, cus = 256;
    int next_lane = 0;
    int64_t next_seq = 0;
    hipEvent_t t0 = nullptr;
    std::map<std::pair<const void*, int>, int> grids;

    // Persistent grid of a kernel: the workgroups that fit on the GPU at once at the kernel's
    // occupancy (a multiple of 8, one share per XCD).
    int grid(const void* fn, int block, size_t lds = 0) {
        const auto key = std::make_pair(fn, block);
        auto it = grids.find(key);
        if (it != grids.end()) return it->second;
        int per_cu = 0;
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds), "occupancy");
        const int g = std::max(8, (std::max(per_cu, 1) * cus) & ~7);
        grids.emplace(key, g);
        return g;
    }

    // Collects lane L's batch: waits for its program (the lane's staging block and scratch are
    // then free; copies of host outputs may still run, finish_all waits for those) and
    // accumulates its times.
    void finish(Lane& L, mct_timing* tm) {
        if (!L.pending) return;
        hip_check(hipEventSynchronize(L.done), "batch");
        L.pending = false;
        if (!L.timed) return;
        L.timed = false;
        for (int k = 0; k < kKCount; ++k) {
            if (!L.kernel_ev[k][0]) continue;
            float ms = 0;
            if (hipEventElapsedTime(&ms, L.kernel_ev[k][0], L.kernel_ev[k][1]) == hipSuccess) {
                kernel_ms[k] += ms;
                kernel_launches[k] += 1;
            }
            (void)hipEventDestroy(L.kernel_ev[k][0]);
            (void)hipEventDestroy(L.kernel_ev[k][1]);
            L.kernel_ev[k][0] = L.kernel_ev[k][1] = nullptr;
        }
        if (!tm) return;
        float t[kStCount];
        for (int i = 0; i < kStCount; ++i) hip_check(hipEventElapsedTime(&t[i], L.stage_ev[i], L.stage_ev[i + 1]), "event time");
        tm->h2d_ms += t[kStH2D];
        tm->biome_ms += t[kStBiome];
        tm->noise_ms += t[kStNoise];
        tm->aquifer_ms += t[kStAquifer];
        tm->fill_ms += t[kStFill];
        tm->surface_ms += t[kStSurface];
        if (L.copies) {
            float c = 0;
            hip_check(hipEventSynchronize(L.copy_ev[1]), "copy");
            hip_check(hipEventElapsedTime(&c, L.copy_ev[0], L.copy_ev[1]), "event time");
            tm->d2h_ms += c;
        }
    }

    static int passes(int bits) { return (bits + 7) / 8; }

    // Position-key codec of a batch's chunks: x / z in units of `unit` per chunk extended by
    // [lo, hi] units, seed slots (make_codec).
    KeyCodec codec(const ChunkRef* ch, int n, int unit, int xlo, int xhi, int zlo, int zhi) const {
        try {
            return make_codec(ch, n, unit, xlo, xhi, zlo, zhi);
        } catch (const std::range_error& e) {
            throw Error(MCT_EINVAL, e.what());
        }
    }

    // Tables of the batch's seeds: resident ones are reused, missing ones get a table (least
    // recently used first) and a build job. Waits for in-flight batches that still read a
    // replaced table; makes lane L wait for batches of other lanes still building a table it
    // reads.
    void plan_seeds(Lane& L, const int64_t* sd, int n, int64_t seq, ChunkRef* ch, SeedJob* jobs, int& njobs) {
        SeedCache& S = seeds;
        njobs = 0;
        std::unordered_map<int64_t, int> batch;   // seed -> table, for this batch
        for (int i = 0; i < n; ++i) {
            auto bt = batch.find(sd[i]);
            if (bt == batch.end()) {
                int t;
                auto it = S.where.find(sd[i]);
                if (it != S.where.end()) {
                    t = it->second;
                    // A lane runs its next batch only after its previous one was collected, so
                    // the table's builder is unfinished only while it is its lane's pending batch.
                    const int bl = S.built_lane[t];
                    if (bl >= 0 && lanes[static_cast<size_t>(bl)].get() != &L) {
                        const Lane& B = *lanes[static_cast<size_t>(bl)];
                        if (B.pending && B.seq == S.built_by[t]) hip_check(hipStreamWaitEvent(L.stream, B.done, 0), "wait for seed tables");
                    }
                } else {
                    t = -1;
                    for (int k = 0; k < S.cap; ++k) {
                        if (S.last_use[k] == seq) continue;   // read by this batch
                        if (t < 0 || S.last_use[k] < S.last_use[t]) t = k;
                    }
                    if (t < 0) throw Error(MCT_ECAPACITY, "more distinct seeds in one batch than seed_capacity");
                    if (S.last_use[t] >= 0) {
                        // the batches that read the table are the pending ones up to its last use
                        for (auto& l : lanes)
                            if (l->pending && l->seq <= S.last_use[t]) finish(*l, nullptr);
                        S.where.erase(S.seed[t]);
                    }
                    S.seed[t] = sd[i];
                    S.where[sd[i]] = t;
                    S.built_by[t] = seq;
                    S.built_lane[t] = lane_index(L);
                    jobs[njobs++] = SeedJob{sd[i], t, 0};
                }
                S.last_use[t] = seq;
                bt = batch.emplace(sd[i], t).first;
            }
            ch[i].seed = bt->second;
        }
    }
    int lane_index(const Lane& L) const {
        for (size_t i = 0; i < lanes.size(); ++i)
            if (lanes[i].get() == &L) return static_cast<int>(i);
        return 0;
    }

    // Stages one batch in lane L's pinned block and launches its program (graph or timed).
    void launch(Lane& L, const int64_t* sd, const int32_t* xz, int n, const BatchOut& out, bool core, bool timed) {
        LaneMemory& m = *L.mem;
        const int64_t seq = next_seq++;
        // the lane's previous batch has been collected (finish), so its staging block is free
        BatchDesc& D = *m.h_desc();
        ChunkRef* ch = m.h_chunks();
        for (int i = 0; i < n; ++i) ch[i] = ChunkRef{xz[2 * i], xz[2 * i + 1], 0, 0};
        int njobs = 0;
        plan_seeds(L, sd, n, seq, ch, m.h_jobs(), njobs);
        D = BatchDesc{};
        D.n = n;
        D.njobs = njobs;
        D.col_keys = n * kCols;
        D.grid_keys = n * kGridKeys;
        D.bio_keys = n * 9;
        D.kc = codec(ch, n, 4, 0, 4, 0, 4);
        D.kg = codec(ch, n, 4, -16, 10, -8, 10);
        D.kb = codec(ch, n, 1, -1, 1, -1, 1);
        Shape sh;
        sh.core = core;
        {   // chunk-count bucket: a multiple of an eighth of the lane's capacity (at least 64)
            const int step = std::max(64, (m.cap + 7) / 8);
            sh.grid_n = std::min(m.cap, (n + step - 1) / step * step);
        }
        const auto pick = [&](const KeyCodec& k) { return dense_size(k) <= m.dense_cap ? kDensePass : passes(k.bits); };
        sh.pass_col = pick(D.kc);
        sh.pass_grid = core ? 0 : pick(D.kg);
        sh.pass_bio = core ? 0 : pick(D.kb);
        // Host outputs: the batch writes result buffer k of the lane (alternating, so the copy
        // of the lane's previous batch can still read the other one) unless aquifer or density
        // outputs are asked for, which have one buffer each (then buffer 0 and both waits).
        const bool to_host = !out.device && (out.terrain || out.biomes || out.aquifer || out.density);
        const bool single = out.aquifer || out.density;
        int k = 0;
        if (to_host && !single) {
            k = L.obuf;
            L.obuf ^= 1;
            if (k == 1) {
                m.ensure(&m.blocks2, static_cast<size_t>(m.cap) * kPer);
                m.ensure(&m.bout2, static_cast<size_t>(m.cap) * kQPer);
            }
        }
        uint8_t* tbuf = k ? m.blocks2 : m.blocks;
        uint8_t* bbuf = k ? m.bout2 : m.bout;
        if (out.device) {
            D.terrain = out.terrain;
            D.biomes = out.biomes;
            D.aquifer = out.aquifer;
            D.density = out.density;
        } else {
            D.terrain = out.terrain ? tbuf : nullptr;
            if (out.biomes) D.biomes = bbuf;
            if (out.aquifer) {
                m.ensure(&m.aq, static_cast<size_t>(m.cap) * kPer);
                D.aquifer = m.aq;
            }
            if (out.density) {
                m.ensure(&m.density, static_cast<size_t>(m.cap) * kPer);
                D.density = m.density;
            }
        }
        if (!core && !D.terrain) D.terrain = tbuf;   // the full step always fills
        sh.biomes = D.biomes != nullptr;
        sh.aquifer = D.aquifer != nullptr;
        sh.density = D.density != nullptr;
        sh.fill = D.terrain != nullptr;
        // the buffers this batch writes are free once the copies that read them are done
        if (to_host && !single) {
            hip_check(hipStreamWaitEvent(L.stream, L.buf_free[k], 0), "wait for copy back");
        } else {
            for (hipEvent_t e : L.buf_free) hip_check(hipStreamWaitEvent(L.stream, e, 0), "wait for copy back");
        }
        L.seq = seq;
        L.timed = timed;
        if (timed) {
            run_program(L, sh, true);
        } else if (use_graphs) {
            auto it = L.graphs.find(sh);
            if (it == L.graphs.end()) it = L.graphs.emplace(sh, capture(L, sh)).first;
            hip_check(hipGraphLaunch(it->second, L.stream), "graph launch");
        } else {
            run_program(L, sh, false);
        }
        hip_check(hipEventRecord(L.done, L.stream), "event");
        L.copies = to_host;
        if (L.copies) {
            if (!copy_stream) hip_check(hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking), "stream");
            hipStream_t cs = copy_stream;
            hip_check(hipStreamWaitEvent(cs, L.done, 0), "wait for batch");
            if (timed) hip_check(hipEventRecord(L.copy_ev[0], cs), "event");
            const size_t nn = static_cast<size_t>(n);
            if (out.terrain) hip_check(hipMemcpyAsync(out.terrain, tbuf, nn * kPer, hipMemcpyDeviceToHost, cs), "copy terrain");
            if (out.biomes) hip_check(hipMemcpyAsync(out.biomes, bbuf, nn * kQPer, hipMemcpyDeviceToHost, cs), "copy biomes");
            if (out.aquifer) hip_check(hipMemcpyAsync(out.aquifer, m.aq, nn * kPer, hipMemcpyDeviceToHost, cs), "copy aquifer");
            if (out.density) hip_check(hipMemcpyAsync(out.density, m.density, nn * kPer * sizeof(float), hipMemcpyDeviceToHost, cs), "copy density");
            if (timed) hip_check(hipEventRecord(L.copy_ev[1], cs), "event");
            record_batch(L, cs);
            hip_check(hipEventRecord(L.copied, cs), "event");
            if (single) {
                for (hipEvent_t e : L.buf_free) hip_check(hipEventRecord(e, cs), "event");
            } else {
                hip_check(hipEventRecord(L.buf_free[k], cs), "event");
            }
        } else {
            record_batch(L, L.stream);
            hip_check(hipEventRecord(L.copied, L.stream), "event");
        }
        L.pending = true;
    }

    // The completion event of the lane's batch in the current call (its results in place),
    // recorded on `s` before the lane's `copied` event, which finish_all waits for.
    void record_batch(const Lane& L, hipStream_t s) {
        if (L.batch_index < 0) return;
        while (batch_ev.size() <= static_cast<size_t>(L.batch_index)) {
            hipEvent_t e = nullptr;
            hip_check(hipEventCreate(&e), "event");
            batch_ev.push_back(e);
        }
        hip_check(hipEventRecord(batch_ev[static_cast<size_t>(L.batch_index)], s), "event");
    }

    hipGraphExec_t capture(Lane& L, const Shape& sh) {
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
            run_program(L, sh, false);
        } catch (...) {
            (void)hipStreamEndCapture(L.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(L.stream, &g), "end capture");
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(e, "graph instantiate");
        return exec;
    }

    // Events around a kernel group (timed batches only).
    template <typename F>
    void group(Lane& L, bool timed, int id, F&& launch) {
        if (!timed) {
            launch();
            return;
        }
        if (!L.kernel_ev[id][0]) {
            hip_check(hipEventCreate(&L.kernel_ev[id][0]), "event");
            hip_check(hipEventCreate(&L.kernel_ev[id][1]), "event");
        }
        hip_check(hipEventRecord(L.kernel_ev[id][0], L.stream), "event");
        launch();
        hip_check(hipEventRecord(L.kernel_ev[id][1], L.stream), "event");
    }
    void stage(Lane& L, bool timed, int s) {
        if (timed) hip_check(hipEventRecord(L.stage_ev[s], L.stream), "event");
    }

    // Dedupe of the keys at `keys` (count at device `np`, codec at device `K`): unique keys
    // ascending, map, count. npass: the radix passes, or kDensePass for the dense form.
    void dedupe(Lane& L, const uint64_t* keys, const int* np, const KeyCodec* K, int npass, int cap, uint64_t* unique, int* map, int* count) {
        LaneMemory& m = *L.mem;
        hipStream_t s = L.stream;
        if (npass == kDensePass) {
            k_dense_clear<<<grid((const void*)k_dense_clear, 256), 256, 0, s>>>(K, m.dense_f);
            k_dense_mark<<<grid((const void*)k_dense_mark, 256), 256, 0, s>>>(keys, np, K, m.dense_f);
            k_dense_tiles<<<grid((const void*)k_dense_tiles, 256), 256, 0, s>>>(K, m.dense_f, m.dense_t);
            k_dense_offsets<<<1, 1024, 0, s>>>(K, m.dense_t, count);
            k_dense_ranks<<<grid((const void*)k_dense_ranks, 256), 256, 0, s>>>(K, m.dense_f, m.dense_t, unique);
            k_dense_map<<<grid((const void*)k_dense_map, 256), 256, 0, s>>>(keys, np, K, m.dense_f, map);
            hip_check(hipGetLastError(), "dense dedupe kernels");
            return;
        }
        const unsigned tiles = blocks_for(static_cast<size_t>(cap), kRsTile);
        const unsigned stiles = blocks_for(256 * static_cast<size_t>(tiles), kScanTile);
        const uint64_t* kin = keys;
        const int* vin = nullptr;
        uint64_t* kout = m.dka;
        int* vout = m.dva;
        for (int p = 0; p < npass; ++p) {
            const int shift = 8 * p;
            k_rs_hist<<<tiles, kRsThreads, 0, s>>>(kin, np, shift, m.dhist);
            k_scan_tiles<<<stiles, 256, 0, s>>>(m.dhist, np, m.dhsum);
            k_scan_small<<<1, 1024, 0, s>>>(m.dhsum, np, 0, nullptr);
            k_scan_apply<<<stiles, 256, 0, s>>>(m.dhist, np, m.dhsum);
            k_rs_scatter<<<tiles, kRsThreads, 0, s>>>(kin, vin, np, shift, m.dhist, kout, vout);
            kin = kout;
            vin = vout;
            kout = kout == m.dka ? m.dkb : m.dka;
            vout = vout == m.dva ? m.dvb : m.dva;
        }
        k_heads<<<tiles * 16, 256, 0, s>>>(kin, np, m.dflag);
        k_tile_sums<<<tiles, 256, 0, s>>>(m.dflag, np, m.dtsum);
        k_scan_small<<<1, 1024, 0, s>>>(m.dtsum, np, 1, m.dtotal);
        k_unique_scatter<<<tiles, 256, 0, s>>>(kin, vin, m.dflag, np, m.dtsum, unique, map, count, m.dtotal);
        hip_check(hipGetLastError(), "dedupe kernels");
    }

    // The launches of one batch program on lane L's stream.
    void run_program(Lane& L, const Shape& sh, bool timed) {
        LaneMemory& m = *L.mem;
        hipStream_t s = L.stream;
        const BatchDesc* B = m.d_desc();
        const ChunkRef* ch = m.d_chunks();
        const SeedDev* sd = seeds.d;
        stage(L, timed, kStH2D);
        hip_check(hipMemcpyAsync(m.d_block, m.h_block, m.block_bytes, hipMemcpyHostToDevice, s), "upload batch");
        group(L, timed, kKSeedSetup, [&] {
            k_seed_layers<<<grid((const void*)k_seed_layers, kSetupThreads), kSetupThreads, 0, s>>>(seeds.d, B, m.d_jobs(), model.nnormal);
            k_seed_serial<<<grid((const void*)k_seed_serial, kSetupThreads), kSetupThreads, 0, s>>>(seeds.d, B, m.d_jobs(), model.blended_first);
            k_seed_yaxes<<<grid((const void*)k_seed_yaxes, 256), 256, 0, s>>>(seeds.d, B, m.d_jobs());
        });
        hip_check(hipGetLastError(), "seed setup kernels");
        const int* cnt = m.counts;
        stage(L, timed, kStBiome);
        if (!sh.core) {
            const int* nbk = cnt + kCntBiomeChunks;
            group(L, timed, kKBiomeUnion, [&] {
                k_bio_keys<<<grid((const void*)k_bio_keys, 256), 256, 0, s>>>(B, ch, m.bkeys);
                dedupe(L, m.bkeys, m.d_count_of(offsetof(BatchDesc, bio_keys)), m.d_codec(offsetof(BatchDesc, kb)), sh.pass_bio, m.cap * 9, m.ubk, m.nb, m.counts + kCntBiomeChunks);
                k_bio_chunks<<<grid((const void*)k_bio_chunks, 256), 256, 0, s>>>(B, m.ubk, nbk, m.bch, d_stats);
            });
            group(L, timed, kKBiomeColumns, [&] {
                k_biome_columns<<<grid((const void*)k_biome_columns, W(kKBiomeColumns)), W(kKBiomeColumns), 0, s>>>(sd, m.bch, nbk, m.bcol);
            });
            group(L, timed, kKBiomeQuarts, [&] {
                k_biome_targets<<<grid((const void*)k_biome_targets, 256), 256, 0, s>>>(sd, m.bch, nbk, m.bcol, m.tgt());
                const int w = W(kKBiomeQuarts);
                const size_t bk_max = 9 * static_cast<size_t>(sh.grid_n);
                if (model.treef) {
                    if (model.vary_groups > 0) {
                        k_biome_groups<<<blocks_for(bk_max * 16 * model.vary_groups, w), w, 0, s>>>(nbk, model.vary_groups, m.tgt(), m.lstar, m.mind, m.gres);
                        k_biome_quart_final<<<grid((const void*)k_biome_quart_final, 256), 256, 0, s>>>(nbk, model.vary_groups, m.tvary, m.gres, m.lstar, m.mind);
                    } else {
                        k_biome_search_cols<<<blocks_for(bk_max * 16, w), w, 0, s>>>(nbk, m.tgt(), m.lstar, m.mind);
                    }
                    k_biome_search_wide<<<grid((const void*)k_biome_search_wide, 256), 256, 0, s>>>(nbk, m.tgt(), m.lstar, m.mind);
                } else {
                    k_biome_search_scalar<<<grid((const void*)k_biome_search_scalar, w), w, 0, s>>>(nbk, m.tgt(), m.lstar, m.mind, run_len);
                }
            });
            group(L, timed, kKBiomeTies, [&] {
                k_biome_ties<<<grid((const void*)k_biome_ties, W(kKBiomeTies)), W(kKBiomeTies), 0, s>>>(nbk, m.tgt(), m.lstar, m.mind, m.tie);
            });
            group(L, timed, kKBiomeChain, [&] {
                k_biome_chain<<<grid((const void*)k_biome_chain, W(kKBiomeChain)), W(kKBiomeChain), 0, s>>>(nbk, m.tgt(), m.lstar, m.mind, m.tie, m.bb, d_stats);
            });
            k_biome_masks<<<grid((const void*)k_biome_masks, 256), 256, 0, s>>>(nbk, m.bb, m.bmask);
            hip_check(hipGetLastError(), "biome kernels");
        }
        // density corners
        stage(L, timed, kStNoise);
        k_col_keys<<<grid((const void*)k_col_keys, 256), 256, 0, s>>>(B, ch, m.ckeys);
        group(L, timed, kKDedupeCols, [&] {
            dedupe(L, m.ckeys, m.d_count_of(offsetof(BatchDesc, col_keys)), m.d_codec(offsetof(BatchDesc, kc)), sh.pass_col, m.cap * kCols, m.ucols, m.colmap, m.counts + kCntColumns);
        });
        group(L, timed, kKColumns, [&] {
            k_columns<<<grid((const void*)k_columns, W(kKColumns)), W(kKColumns), 0, s>>>(B, sd, m.ucols, cnt + kCntColumns, m.colv);
        });
        group(L, timed, kKCorners, [&] {
            k_corners<<<blocks_for(static_cast<size_t>(sh.grid_n) * kCols, kCornerCols), 256, 0, s>>>(B, sd, m.ucols, cnt + kCntColumns, m.colv, m.corners);
        });
        hip_check(hipGetLastError(), "noise kernels");
        stage(L, timed, kStAquifer);
        if (sh.core) {
            stage(L, timed, kStFill);
            FillArgs F{sd, ch, m.corners, m.colmap, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
            group(L, timed, kKFill, [&] { k_fill_core<<<grid((const void*)k_fill_core, W(kKFill)), W(kKFill), 0, s>>>(F, B); });
            hip_check(hipGetLastError(), "fill kernel");
            stage(L, timed, kStSurface);
            stage(L, timed, kStCount);
            return;
        }
        // surface-level grid and aquifer cells
        k_grid_keys<<<grid((const void*)k_grid_keys, 256), 256, 0, s>>>(B, ch, m.gkeys);
        group(L, timed, kKDedupeGrid, [&] {
            dedupe(L, m.gkeys, m.d_count_of(offsetof(BatchDesc, grid_keys)), m.d_codec(offsetof(BatchDesc, kg)), sh.pass_grid, m.cap * kGridKeys, m.ug, m.gmap, m.counts + kCntGrid);
        });
        const size_t gn = static_cast<size_t>(sh.grid_n);
        group(L, timed, kKFts, [&] { k_fts<<<blocks_for(gn * kGridKeys, W(kKFts)), W(kKFts), 0, s>>>(B, sd, m.ug, cnt + kCntGrid, m.gv); });
        group(L, timed, kKGridScatter, [&] {
            k_grid_scatter<<<grid((const void*)k_grid_scatter, 256), 256, 0, s>>>(B, m.gmap, m.gv, m.grid, m.prelim);
            k_aq_skip<<<grid((const void*)k_aq_skip, 256), 256, 0, s>>>(B, m.grid, m.skip);
        });
        group(L, timed, kKAqCells, [&] {
            k_aq_cells<<<blocks_for(gn * kAqCells, W(kKAqCells)), W(kKAqCells), 0, s>>>(sd, B, ch, m.grid, m.cloc, m.cfl);
        });
        hip_check(hipGetLastError(), "aquifer kernels");
        // fill
        stage(L, timed, kStFill);
        FillArgs F{sd, ch, m.corners, m.colmap, m.cloc, m.cfl, m.skip, nullptr, nullptr, nullptr, nullptr};
        FillArgs FH = F;
        FH.h0 = m.h0;
        FH.clsw = m.clsw;
        group(L, timed, kKHeights, [&] { k_heights_init<<<grid((const void*)k_heights_init, W(kKHeights)), W(kKHeights), 0, s>>>(B, m.h0, m.counts + kCntRedo); });
        group(L, timed, kKFill, [&] {
            const size_t lds = fill_lds_bytes(W(kKFill));
            k_fill<<<static_cast<unsigned>(gn * 256 * kCellsY / static_cast<size_t>(W(kKFill))), W(kKFill), lds, s>>>(FH, B);
        });
        hip_check(hipGetLastError(), "fill kernel");
        if (sh.aquifer) k_aquifer_copy<<<grid((const void*)k_aquifer_copy, 256), 256, 0, s>>>(B);
        // surface
        stage(L, timed, kStSurface);
        SurfArgs A{sd, nullptr, ch, nullptr, m.nb, m.bb, m.corners, m.colmap, m.prelim, m.colrec, m.ore, m.bmask, m.clsw};
        if (kOreRows > 0) group(L, timed, kKOre, [&] { k_ore<<<blocks_for(gn * 256, W(kKOre)), W(kKOre), 0, s>>>(A, B); });
        group(L, timed, kKSurfaceCols, [&] {
            k_surface_cols<<<blocks_for(gn * 256, W(kKSurfaceCols)), W(kKSurfaceCols), 0, s>>>(sd, B, ch, m.colrec, m.nb, m.bmask);
        });
        group(L, timed, kKSurface1, [&] { k_surface1<<<blocks_for(gn * 256, 64), 64, 0, s>>>(A, B, m.h0, m.h1, m.h2); });
        group(L, timed, kKSurface2, [&] { k_surface2<<<static_cast<unsigned>(gn * 4), 64, 0, s>>>(A, B, m.h0, m.h1, m.h2, m.hf); });
        group(L, timed, kKSurfaceExact, [&] {
            k_surface_check<<<grid((const void*)k_surface_check, 256), 256, 0, s>>>(B, m.h2, m.hf, m.flags, m.counts + kCntRedo, d_stats);
            k_surface_exact<<<64, 256, 0, s>>>(F, A, B, m.h0, m.h1, m.h2, m.hf, m.flags, m.counts + kCntRedo, replay_all ? 1 : 0);
        });
        if (sh.biomes) group(L, timed, kKGather, [&] { k_gather_biomes<<<grid((const void*)k_gather_biomes, 256), 256, 0, s>>>(B, m.nb, m.bb); });
        hip_check(hipGetLastError(), "surface kernels");
        stage(L, timed, kStCount);
    }

public:
    // ---- biome-only queries --------------------------------------------------------------------
    // In groups of at most kQueryGroup blocks and seed_cap distinct seeds: local seed ids through
    // a small open-addressing table, queries ordered by seed (counting sort) so that a wave's
    // lanes share a seed's tables, answers scattered back to input order. The tables of a
    // group's seeds are built by a batch of no chunks on lane 0 first.
    static constexpr int kQueryGroup = 1 << 20;
    int32_t *d_qxyz = nullptr, *d_qslot = nullptr, *d_qbad = nullptr;
    uint8_t* d_qout = nullptr;
    void biome_query(const int64_t* sd, const int32_t* xyz, int count, uint8_t* out) {
        if (!model.treef) throw Error(MCT_ECAPACITY, "biome queries need the float climate tree (a bound of this model exceeds 2^23)");
        finish_all(nullptr);
        Lane& L = *lanes[0];
        if (!d_qxyz) {
            hip_check(hipMalloc(&d_qxyz, sizeof(int32_t) * 3 * kQueryGroup), "malloc");
            hip_check(hipMalloc(&d_qslot, sizeof(int32_t) * kQueryGroup), "malloc");
            hip_check(hipMalloc(&d_qout, kQueryGroup), "malloc");
            hip_check(hipMalloc(&d_qbad, sizeof(int)), "malloc");
        }
        int tsize = 16;
        while (tsize < 4 * seed_cap) tsize <<= 1;
        std::vector<int64_t> tkey(static_cast<size_t>(tsize));
        std::vector<int32_t> tval(static_cast<size_t>(tsize));
        std::vector<int32_t> lid, order, slot, cnt, sxyz;
        std::vector<uint8_t> sout;
        for (int off = 0; off < count;) {
            int n = std::min(kQueryGroup, count - off);
            std::fill(tval.begin(), tval.end(), -1);
            std::vector<int64_t> distinct;
            lid.resize(static_cast<size_t>(n));
            for (int i = 0; i < n; ++i) {
                const int64_t s = sd[off + i];
                const uint64_t h = static_cast<uint64_t>(s) * 0x9E3779B97F4A7C15ull;
                size_t at = static_cast<size_t>(h >> 40) & static_cast<size_t>(tsize - 1);
                while (tval[at] >= 0 && tkey[at] != s) at = (at + 1) & static_cast<size_t>(tsize - 1);
                if (tval[at] < 0) {
                    if (static_cast<int>(distinct.size()) == seed_cap) { n = i; break; }
                    tkey[at] = s;
                    tval[at] = static_cast<int32_t>(distinct.size());
                    distinct.push_back(s);
                }
                lid[i] = tval[at];
            }
            build_tables(L, distinct);
            const int nd = static_cast<int>(distinct.size());
            cnt.assign(static_cast<size_t>(nd) + 1, 0);
            for (int i = 0; i < n; ++i) ++cnt[lid[i] + 1];
            for (int d = 0; d < nd; ++d) cnt[d + 1] += cnt[d];
            order.resize(static_cast<size_t>(n));
            slot.resize(static_cast<size_t>(n));
            sxyz.resize(static_cast<size_t>(n) * 3);
            for (int i = 0; i < n; ++i) order[cnt[lid[i]]++] = i;
            std::vector<int32_t> dslot(static_cast<size_t>(nd));
            for (int d = 0; d < nd; ++d) dslot[d] = seeds.where.at(distinct[d]);
            for (int k = 0; k < n; ++k) {
                const size_t i = static_cast<size_t>(off) + static_cast<size_t>(order[k]);
                slot[k] = dslot[lid[order[k]]];
                sxyz[3 * static_cast<size_t>(k)] = xyz[3 * i];
                sxyz[3 * static_cast<size_t>(k) + 1] = xyz[3 * i + 1];
                sxyz[3 * static_cast<size_t>(k) + 2] = xyz[3 * i + 2];
            }
            hip_check(hipMemcpyAsync(d_qxyz, sxyz.data(), sizeof(int32_t) * 3 * n, hipMemcpyHostToDevice, L.stream), "upload");
            hip_check(hipMemcpyAsync(d_qslot, slot.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, L.stream), "upload");
            hip_check(hipMemsetAsync(d_qbad, 0, sizeof(int), L.stream), "memset");
            k_biome_query<<<blocks_for(static_cast<size_t>(n), 256), 256, 0, L.stream>>>(seeds.d, d_qslot, d_qxyz, n, d_qout, d_qbad);
            hip_check(hipGetLastError(), "query kernel");
            int bad = 0;
            sout.resize(static_cast<size_t>(n));
            hip_check(hipMemcpyAsync(sout.data(), d_qout, static_cast<size_t>(n), hipMemcpyDeviceToHost, L.stream), "download");
            hip_check(hipMemcpyAsync(&bad, d_qbad, sizeof(int), hipMemcpyDeviceToHost, L.stream), "download");
            hip_check(hipStreamSynchronize(L.stream), "sync");
            if (bad) throw Error(MCT_ECAPACITY, std::to_string(bad) + " queries have climate targets beyond 2^23 (not served by the float tree)");
            for (int k = 0; k < n; ++k) out[off + order[k]] = sout[k];
            off += n;
        }
    }

private:
    // Makes the given seeds' tables resident: the seed-setup launches of a batch of no chunks.
    void build_tables(Lane& L, const std::vector<int64_t>& need) {
        LaneMemory& m = *L.mem;
        finish(L, nullptr);
        const int64_t seq = next_seq++;
        std::vector<ChunkRef> tmp(need.size());
        int njobs = 0;
        plan_seeds(L, need.data(), static_cast<int>(need.size()), seq, tmp.data(), m.h_jobs(), njobs);
        L.seq = seq;
        if (njobs == 0) return;
        BatchDesc& D = *m.h_desc();
        D = BatchDesc{};
        D.njobs = njobs;
        hip_check(hipStreamWaitEvent(L.stream, L.copied, 0), "wait");
        hip_check(hipMemcpyAsync(m.d_block, m.h_block, m.block_bytes, hipMemcpyHostToDevice, L.stream), "upload jobs");
        k_seed_layers<<<grid((const void*)k_seed_layers, kSetupThreads), kSetupThreads, 0, L.stream>>>(seeds.d, m.d_desc(), m.d_jobs(), model.nnormal);
        k_seed_serial<<<grid((const void*)k_seed_serial, kSetupThreads), kSetupThreads, 0, L.stream>>>(seeds.d, m.d_desc(), m.d_jobs(), model.blended_first);
        k_seed_yaxes<<<grid((const void*)k_seed_yaxes, 256), 256, 0, L.stream>>>(seeds.d, m.d_desc(), m.d_jobs());
        hip_check(hipGetLastError(), "seed setup kernels");
        hip_check(hipEventRecord(L.done, L.stream), "event");
        hip_check(hipEventRecord(L.copied, L.stream), "event");
        L.copies = false;
        L.batch_index = -1;
        L.pending = true;
        finish(L, nullptr);
    }
};

}  // namespace mct
