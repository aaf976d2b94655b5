// gfx950 TERRAIN-step generator: kernels, device radix-sort dedupe, per-seed setup on the
// device, batch pipeline on HIP streams, and the C ABI declared in include/mct.h.
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
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mct.h"
#include "pipeline.h"

using namespace mct;

namespace {

__constant__ ModelDev c_model;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(MCT_EDEVICE, std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// ---- per-seed setup kernels -------------------------------------------------------------
struct SeedJob {
    int64_t seed;
    int32_t slot;
    int32_t pad;
};

constexpr int kSetupThreads = 64;

// One thread per (job, NormalNoise layer); the permutation is shuffled in the thread's LDS row.
__global__ __launch_bounds__(kSetupThreads) void k_seed_layers(SeedDev* seeds, const SeedJob* jobs, int njobs, int nnormal) {
    __shared__ uint8_t scratch[kSetupThreads][256];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= njobs * nnormal) return;
    const SeedJob j = jobs[idx / nnormal];
    seed_layer(seeds[j.slot], j.seed, idx % nnormal, scratch[threadIdx.x]);
}

// One thread per (job, BlendedNoise node): 40 octaves in sequence; one thread per job for
// the positional factories, clay bands and zoom seed.
__global__ __launch_bounds__(kSetupThreads) void k_seed_serial(SeedDev* seeds, const SeedJob* jobs, int njobs, const int* blended_first) {
    __shared__ uint8_t scratch[kSetupThreads][256];
    const int per = mctgen::kNumBlended + 1;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= njobs * per) return;
    const SeedJob j = jobs[idx / per];
    const int k = idx % per;
    if (k < mctgen::kNumBlended) seed_blended(seeds[j.slot], j.seed, k, blended_first[k], scratch[threadIdx.x]);
    else seed_misc(seeds[j.slot], j.seed, c_model);
}

// ---- biomes ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_biome_columns(const SeedDev* seeds, const ChunkRef* bch, int nbk, float* bcol) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nbk * 16) biome_column(seeds, bch, bcol, idx);
}

__global__ __launch_bounds__(256) void k_biome_targets(const SeedDev* seeds, const ChunkRef* bch, int nbk, const float* bcol, int64_t* targets) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nbk * kQPer) biome_target(seeds, bch, bcol, targets, idx);
}

// R-tree searches, one run of kBiomeRun quarts per thread; the upper levels of the tree (the
// first kTreeTop nodes of the breadth-first order, 24 KB) are staged in LDS once per
// workgroup. Quarts with targets beyond the float tree's range are left to k_biome_search_wide.
constexpr int kTreeTop = 384;
constexpr int kSearchDepth = 12;
__global__ __launch_bounds__(256) void k_biome_search(int nbk, const int64_t* targets, int32_t* lstar, int64_t* mind, int run_len) {
    __shared__ TreeNodeF top[kTreeTop];
    const ModelDev& M = c_model;
    const int ntop = mctgen::kTreeNodes < kTreeTop ? mctgen::kTreeNodes : kTreeTop;
    for (int i = threadIdx.x; i < ntop * 16; i += blockDim.x) reinterpret_cast<uint32_t*>(top)[i] = reinterpret_cast<const uint32_t*>(M.treef)[i];
    __syncthreads();
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nbk * 16 * biome_runs(run_len)) biome_search(M, top, ntop, targets, lstar, mind, idx, false, run_len);
}

// L* for whole quart columns (one thread per column): a restricted search per group of the
// column's varying coordinate instead of one search per quart.
__global__ __launch_bounds__(256) void k_biome_search_cols(int nbk, const int64_t* targets, int32_t* lstar, int64_t* mind) {
    __shared__ TreeNodeF top[kTreeTop];
    const ModelDev& M = c_model;
    const int ntop = mctgen::kTreeNodes < kTreeTop ? mctgen::kTreeNodes : kTreeTop;
    for (int i = threadIdx.x; i < ntop * 16; i += blockDim.x) reinterpret_cast<uint32_t*>(top)[i] = reinterpret_cast<const uint32_t*>(M.treef)[i];
    __syncthreads();
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nbk * 16) biome_search_column(M, top, ntop, targets, lstar, mind, idx);
}

// The column search split in two: group searches (thread per column and group), then each
// quart's L* (thread per quart).
__global__ __launch_bounds__(256) void k_biome_groups(int nbk, int G, const int64_t* targets, int32_t* lstar, int64_t* mind, GroupRes* res) {
    __shared__ TreeNodeF top[kTreeTop];
    const ModelDev& M = c_model;
    const int ntop = mctgen::kTreeNodes < kTreeTop ? mctgen::kTreeNodes : kTreeTop;
    for (int i = threadIdx.x; i < ntop * 16; i += blockDim.x) reinterpret_cast<uint32_t*>(top)[i] = reinterpret_cast<const uint32_t*>(M.treef)[i];
    __syncthreads();
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nbk * 16 * G) biome_group_search(M, top, ntop, targets, lstar, mind, res, G, idx);
}

__global__ __launch_bounds__(256) void k_biome_quart_final(int nbk, int G, const int64_t* targets, const GroupRes* res, int32_t* lstar, int64_t* mind) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nbk * kQPer) biome_quart_final(c_model, targets, res, lstar, mind, G, idx);
}

__global__ __launch_bounds__(256) void k_biome_search_wide(int nbk, const int64_t* targets, int32_t* lstar, int64_t* mind) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nbk * kQPer) biome_search_wide(c_model, targets, lstar, mind, idx);
}

// The search without the float tree (a model whose bounds it cannot hold exactly).
__global__ __launch_bounds__(256) void k_biome_search_scalar(int nbk, const int64_t* targets, int32_t* lstar, int64_t* mind, int run_len) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nbk * 16 * biome_runs(run_len)) biome_search(c_model, nullptr, 0, targets, lstar, mind, idx, true, run_len);
}

__global__ __launch_bounds__(256) void k_biome_ties(int nbk, const int64_t* targets, const int32_t* lstar, const int64_t* mind, uint8_t* tie) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nbk * kQPer) biome_tie(c_model, targets, lstar, mind, tie, idx);
}

// The last-result rule, one wave64 per biome chunk. Windows of 64 chain positions with no tie
// flag, entered in sync (previous result == previous L*), resolve to L* in parallel; any other
// window is walked by lane 0 with the full rule.
__global__ __launch_bounds__(256) void k_biome_chain(int nbk, const int64_t* targets, const int32_t* lstar, const int64_t* mind, const uint8_t* tie,
                                                     uint8_t* biomes, unsigned long long* stats) {
    // One wave per biome chunk, 64 chain positions per step. While the chain is in sync (the
    // previous result is the previous position's L*) a position's result is its L* unless its
    // tie flag is set, where the previous L* wins; from there the chain is out of sync with that
    // leaf X, which keeps winning while it is at minimal distance (all lanes test it at once)
    // until it loses or a position's L* is X itself. Each switch costs one ballot.
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= nbk) return;
    const ModelDev& M = c_model;
    const size_t base = static_cast<size_t>(wave) * kQPer;
    int prev = -1;       // result at the previous chain position
    bool sync = true;    // prev == L* of the previous position
    unsigned long long changed = 0, changed_biome = 0;
    for (int p0 = 0; p0 < kQPer; p0 += 64) {
        const int p = p0 + lane;
        const size_t i = base + chain_quart(p);
        const int ls = lstar[i];
        const bool t = p > 0 && tie[base + p];
        int res = ls;
        int k = 0;
        while (k < 64) {
            if (sync) {
                const unsigned long long m = __ballot(t && lane >= k);
                if (m == 0) break;
                const int j = __builtin_ctzll(m);
                const int before = __shfl(ls, j > 0 ? j - 1 : 0);
                const int X = j > 0 ? before : prev;
                if (lane == j) res = X;
                prev = X;
                sync = false;
                k = j + 1;
            } else {
                const bool wins = lane >= k && box_distance(M.tree[M.leaf_node[prev]], targets + i * 7) == mind[i];
                const unsigned long long m = __ballot(lane >= k && (!wins || ls == prev));
                if (m == 0) {
                    if (lane >= k) res = prev;
                    k = 64;
                    break;
                }
                const int j = __builtin_ctzll(m);
                if (lane >= k && lane < j) res = prev;
                prev = __shfl(ls, j);
                sync = true;
                k = j + 1;
            }
        }
        if (sync) prev = __shfl(ls, 63);
        biomes[i] = M.leaf_biome[res];
        changed += static_cast<unsigned long long>(__popcll(__ballot(res != ls)));
        changed_biome += static_cast<unsigned long long>(__popcll(__ballot(res != ls && M.leaf_biome[res] != M.leaf_biome[ls])));
    }
    if (lane == 0 && changed) {
        atomicAdd(stats + 1, changed);
        atomicAdd(stats + 2, changed_biome);
    }
}

// The biomes of each quart column of a biome chunk (a superset of any block's zoom candidates
// taken from it), so the surface rules can skip the zoom where no candidate can matter.
__global__ __launch_bounds__(256) void k_biome_masks(int nbk, const uint8_t* bb, uint64_t* mask) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nbk * 16) return;
    const uint8_t* q = bb + static_cast<size_t>(idx >> 4) * kQPer + (idx & 15);
    uint64_t m0 = 0, m1 = 0;
    for (int qy = 0; qy < kQH; ++qy) {
        const int b = q[qy * 16];
        if (b < 64) m0 |= 1ull << b;
        else m1 |= 1ull << (b - 64);
    }
    mask[idx * 2] = m0;
    mask[idx * 2 + 1] = m1;
}

// ---- density ------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_col_keys(KeyCodec K, const ChunkRef* ch, int n, uint64_t* keys) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n * kCols) column_key(K, ch, keys, idx);
}

__global__ __launch_bounds__(256) void k_columns(KeyCodec K, const SeedDev* seeds, int upper, const uint64_t* ucols, const int* count, float* colv) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < upper && u < *count) column_values(K, seeds, ucols, colv, u);
}

// Corner pass: each workgroup takes kCornerCols distinct columns x kRows rows (245 of 256
// threads). The x / z records of every octave sample site of its columns are computed once
// into LDS (the FP64 coordinate setup of two of the three axes, shared by the 49 rows), then
// each thread evaluates its corner with only the y axis computed per sample. Threads take the
// corners row by row (thread = row * kCornerCols + column), so a wave holds about 13 adjacent
// rows: the terrain's y-dependent branches (cave noises below the surface, entrances above,
// the ore channels' y range) then mostly agree across a wave.
constexpr int kCornerCols = 256 / kRows;
constexpr int kCornerTab = mctgen::kCornerSites > 0 ? mctgen::kCornerSites : 1;
__global__ __launch_bounds__(256) void k_corners(KeyCodec K, const SeedDev* seeds, int upper, const uint64_t* ucols, const int* count, const float* colv,
                                                 float* corners) {
    __shared__ mct::AxisXZ tab[kCornerCols * kCornerTab];
    const int c0 = blockIdx.x * kCornerCols;
    const int ncol = min(min(kCornerCols, upper - c0), *count - c0);   // the same in every thread
    if (ncol <= 0) return;
    for (int k = threadIdx.x; k < ncol * mctgen::kCornerSites; k += blockDim.x) {
        const int cl = k / mctgen::kCornerSites;
        tab[k] = corner_site_axes(K, seeds, ucols, c0 + cl, k - cl * mctgen::kCornerSites);
    }
    __syncthreads();
    const int row = static_cast<int>(threadIdx.x) / kCornerCols, cl = static_cast<int>(threadIdx.x) - row * kCornerCols;
    if (row < kRows && cl < ncol)
        corner_values(K, seeds, ucols, colv, corners, c0 + cl, row, (const MCT_LDS mct::AxisXZ*)(tab + cl * kCornerTab));
}

// ---- surface-level grid / aquifer -----------------------------------------------------------
__global__ __launch_bounds__(256) void k_grid_keys(KeyCodec K, const ChunkRef* ch, int n, uint64_t* keys) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n * kGridKeys) grid_key(K, ch, keys, idx);
}

__global__ __launch_bounds__(256) void k_fts(KeyCodec K, const SeedDev* seeds, int upper, const uint64_t* ukeys, const int* count, float* vals) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < upper && u < *count) fts_value(K, seeds, ukeys, vals, u);
}

__global__ __launch_bounds__(256) void k_grid_scatter(int n, const int* map, const float* vals, int32_t* grid, float* prelim) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n * kGridKeys) grid_scatter(map, vals, grid, prelim, idx);
}

__global__ __launch_bounds__(256) void k_aq_skip(int n, const int32_t* grid, int* skip) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) aquifer_skip(grid, skip, c);
}

__global__ __launch_bounds__(256) void k_aq_cells(const SeedDev* seeds, const ChunkRef* ch, int n, const int32_t* grid, int32_t* loc, int32_t* fl) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n * kAqCells) aquifer_cell(seeds, ch, grid, loc, fl, idx);
}

// ---- fill --------------------------------------------------------------------------------
// A workgroup fills one cell row of one chunk (W threads: W / 16 z rows of 16 x, one block
// column of the row's 8 blocks each); rows are taken from the top down within a chunk, so the
// columns' heights are usually final before the lower rows would raise them (their atomics are
// then skipped). The chunk's aquifer cells the row can read are staged in LDS. The blocks
// whose aquifer barrier / pressure part decides (a few percent of the blocks, but spread over
// most waves that reach the aquifer) are queued in LDS and resolved after the row, densely, by
// all threads of the workgroup. Each column's 8 block classes (2 bits each) go to the class
// words the surface walk reads (SurfArgs::clsw), half a word per cell row.
constexpr int kFillStage = 36;   // 3 x 3 aquifer cells by at most 4 grid rows (8 block rows +-1)
constexpr int kFillHead = kFillStage * 6 + 4;   // ints before the per-thread arrays
static_assert(kFillHead * 4 % 16 == 0, "per-thread LDS arrays stay 16-byte aligned");
constexpr size_t fill_lds_bytes(int w) { return static_cast<size_t>(kFillHead) * 4 + 56 * static_cast<size_t>(w); }

// A staged aquifer cell in chunk-local coordinates (l = location - 16 * chunk in x and z):
// s = |l|^2 and the factors -2 l of the cross terms, so that |l - b|^2 - |b|^2 of a block b of
// the chunk is s + ax * bx + ay * by + az * bz.
struct alignas(16) AqCellTerm {
    int s, ax, ay, az;
};

struct AqCellsLds {
    const MCT_LDS AqCellTerm* loc;   // one per staged cell
    const MCT_LDS int* fl;           // 2 ints per staged cell
    int base;                        // chunk cell index of staged cell 0
    __device__ MCT_FI Fluid fluid(int i) const { return Fluid{fl[(i - base) * 2], fl[(i - base) * 2 + 1]}; }
};

// a * b + c for a, b within +-2^23, as one full-rate v_mad_i32_i24 (the compiler falls back to
// the quarter-rate 32-bit multiply where one factor is wave-uniform).
__device__ MCT_FI int mad_i24(int a, int b, int c) {
    int r;
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

// aquifer_nearest over staged cells. Each distance is |l - b|^2 - |b|^2: the block's own |b|^2
// is common to all 12 cells, so the order, the ties and every difference of two distances are
// those of the plain scan, and only differences are read (d2 - d1, d3 - d1, d3 - d2). Every
// factor is within +-2^10 (cells at chunk-local x, z in -16..25 and y in -84..332), so the products are
// exact 24-bit multiply-adds.
__device__ MCT_FI AqNearest aquifer_nearest(const AqCellsLds& C, int cx, int cz, int x, int y, int z) {
    const int xa = (x - 5) >> 4, ya = grid_yf(y + 1), za = (z - 5) >> 4;
    const int i0 = ((ya - 1 - kAqMinGY) * 3 + (za - (cz - 1))) * 3 + (xa - (cx - 1));
    const int bx = x - cx * 16, bz = z - cz * 16;
    const MCT_LDS AqCellTerm* p = C.loc + (i0 - C.base);
    int d1 = INT32_MAX, d2 = INT32_MAX, d3 = INT32_MAX, d4 = INT32_MAX;
    int o1 = 0, o2 = 0, o3 = 0;
#pragma unroll
    for (int ox = 0; ox <= 1; ++ox)
#pragma unroll
        for (int oy = 0; oy <= 2; ++oy)
#pragma unroll
            for (int oz = 0; oz <= 1; ++oz) {
                const int o = oy * 9 + oz * 3 + ox;
                const AqCellTerm t = p[o];
                const int nd = mad_i24(t.az, bz, mad_i24(t.ay, y, mad_i24(t.ax, bx, t.s)));
                if (d1 >= nd) { o3 = o2; o2 = o1; o1 = o; d4 = d3; d3 = d2; d2 = d1; d1 = nd; }
                else if (d2 >= nd) { o3 = o2; o2 = o; d4 = d3; d3 = d2; d2 = nd; }
                else if (d3 >= nd) { o3 = o; d4 = d3; d3 = nd; }
                else if (d4 >= nd) { d4 = nd; }
            }
    return AqNearest{i0 + o1, i0 + o2, i0 + o3, d1, d2, d3};
}

// The 12 cells of the nearest-3 scan for the blocks of one column that share grid row ya
// (y + 1 in [12 ya, 12 ya + 11]), as terms of the scan key (below): per cell 32 times the
// column part s + ax * bx + az * bz of the expanded distance, plus 31 - k, and 32 ay, so that
// a block's key is one multiply-add, ay * y + p (|32 ay| < 2^15, |p| < 2^23). ya is the same
// for all threads of a workgroup (one cell row of one chunk), so reloads are uniform.
struct AqColumnRow {
    AqCellsLds C;
    int i0 = 0, ya = INT32_MIN;
    int p[12], ay[12];
    __device__ MCT_FI void load(int cx, int cz, int x, int z, int y) {
        ya = grid_yf(y + 1);
        const int xa = (x - 5) >> 4, za = (z - 5) >> 4;
        i0 = ((ya - 1 - kAqMinGY) * 3 + (za - (cz - 1))) * 3 + (xa - (cx - 1));
        const int bx = x - cx * 16, bz = z - cz * 16;
        const MCT_LDS AqCellTerm* q = C.loc + (i0 - C.base);
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            const AqCellTerm t = q[(k % 6 >> 1) * 9 + (k & 1) * 3 + k / 6];
            p[k] = mad_i24(t.az, bz, mad_i24(t.ax, bx, t.s)) * 32 + (31 - k);
            ay[k] = t.ay * 32;
        }
    }
    __device__ MCT_FI Fluid fluid(int i) const { return C.fluid(i); }
};

// aquifer_nearest for a block of the loaded column row: the scan of the staged-cell form in
// the same order k = ox * 6 + oy * 2 + oz, each cell ranked by its key 32 * distance + 31 - k.
// The keys are distinct and ordered by distance and, among equal distances, later cells first:
// the order the scan's insertion (d >= nd puts the new cell first) builds. The three smallest
// keys are kept with one min and two med3; distances and cell offsets are decoded at the end.
constexpr uint64_t kAqScanOffsets = [] {
    uint64_t t = 0;
    for (int k = 0; k < 12; ++k) t |= static_cast<uint64_t>((k % 6 >> 1) * 9 + (k & 1) * 3 + k / 6) << (5 * k);
    return t;
}();
__device__ MCT_FI int aq_key_offset(int key) { return static_cast<int>((kAqScanOffsets >> (5 * (31 - (key & 31)))) & 31u); }
__device__ MCT_FI int med3_i32(int a, int b, int c) {
    int r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
__device__ MCT_FI AqNearest aquifer_nearest(const AqColumnRow& R, int, int, int, int y, int) {
    int k1 = INT32_MAX, k2 = INT32_MAX, k3 = INT32_MAX;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        const int key = mad_i24(R.ay[k], y, R.p[k]);
        k3 = med3_i32(k2, key, k3);   // k1 < k2 < k3: the new third is the median of k2, key, k3
        k2 = med3_i32(k1, key, k2);
        k1 = min(k1, key);
    }
    return AqNearest{R.i0 + aq_key_offset(k1), R.i0 + aq_key_offset(k2), R.i0 + aq_key_offset(k3), k1 >> 5, k2 >> 5, k3 >> 5};
}

__global__ __launch_bounds__(256) void k_fill(FillArgs A, int n) {
    extern __shared__ __attribute__((aligned(16))) int fill_lds[];
    const int W = static_cast<int>(blockDim.x), tid = static_cast<int>(threadIdx.x);
    const size_t idx0 = static_cast<size_t>(blockIdx.x) * static_cast<size_t>(W);
    if (idx0 >= static_cast<size_t>(n) * 256 * kCellsY) return;   // whole workgroups only
    const int c = static_cast<int>(idx0 / (256 * kCellsY));
    const int t0 = static_cast<int>(idx0 % (256 * kCellsY));
    const int cy = kCellsY - 1 - (t0 >> 8);
    const int col0 = t0 & 255;
    int* sloc = fill_lds;
    int* sfl = fill_lds + kFillStage * 4;
    int* qn = fill_lds + kFillStage * 6;
    int* hm = fill_lds + kFillHead;
    uint32_t* fr = reinterpret_cast<uint32_t*>(fill_lds + kFillHead + W);   // class bits of the column's 8 blocks
    float* qd = reinterpret_cast<float*>(fill_lds + kFillHead + 2 * W);
    uint16_t* qk = reinterpret_cast<uint16_t*>(fill_lds + kFillHead + 10 * W);
    const ChunkRef cr = A.chunks[c];
    const int ylo = kMinY + cy * 8;
    const int skip = A.skip[c];
    const int gy_lo = max(grid_yf(ylo + 1) - 1, kAqMinGY), gy_hi = min(grid_yf(ylo + 8) + 1, kAqMaxGY);
    const int base = (gy_lo - kAqMinGY) * 9;
    if (ylo <= skip) {   // some block of the row can reach the aquifer cells
        const int32_t* loc = A.cell_loc + static_cast<size_t>(c) * kAqCells * 4;
        const int32_t* fl = A.cell_fluid + static_cast<size_t>(c) * kAqCells * 2;
        for (int k = tid; k < (gy_hi - gy_lo + 1) * 9; k += W) {
            const int lx = loc[(base + k) * 4] - cr.x * 16, ly = loc[(base + k) * 4 + 1], lz = loc[(base + k) * 4 + 2] - cr.z * 16;
            sloc[k * 4] = lx * lx + ly * ly + lz * lz;
            sloc[k * 4 + 1] = -2 * lx;
            sloc[k * 4 + 2] = -2 * ly;
            sloc[k * 4 + 3] = -2 * lz;
            sfl[k * 2] = fl[(base + k) * 2];
            sfl[k * 2 + 1] = fl[(base + k) * 2 + 1];
        }
    }
    if (tid == 0) *qn = 0;
    __syncthreads();
    const AqCellsLds C{(const MCT_LDS AqCellTerm*)sloc, (const MCT_LDS int*)sfl, base};
    AqColumnRow R{C};
    const Octave* O = A.seeds[cr.seed].oct;
    const int x = (col0 + tid) & 15, z = (col0 + tid) >> 4;
    constexpr int NC = mctgen::kNumChannels > 0 ? mctgen::kNumChannels : 1;
    float val[NC], step[NC];
    cell_channels(A, c, x, z, cy, val, step);
    const int bx = cr.x * 16 + x, bz = cr.z * 16 + z;
    uint8_t* out = A.blocks + static_cast<size_t>(c) * kPer + static_cast<size_t>(cy * 8 * 256 + z * 16 + x);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));   // lane id
    int top = -1;
    uint32_t frag = 0;
    int qwave = 0;
    const bool one_wave = W <= 64;   // then the queue needs no LDS atomic
#pragma unroll
    for (int yi = 0; yi < 8; ++yi) {
        const int by = ylo + yi;
        const float d = mctgen::eval_block(O, bx, by, bz, val);
        if (ylo <= skip && grid_yf(by + 1) != R.ya) R.load(cr.x, cr.z, bx, bz, by);
        AqNearest nr;
        const int s = aquifer_front(R, cr.x, cr.z, skip, bx, by, bz, static_cast<double>(d), nr);
        const bool pend = s == kAqPending;
        const uint64_t m = __builtin_amdgcn_ballot_w64(pend);
        if (m) {
            const int lead = __builtin_ctzll(m);
            int slot = qwave;   // one wave per workgroup: the queue length is wave-uniform
            if (!one_wave) {
                if (static_cast<int>(below) == lead) slot = atomicAdd(qn, __builtin_popcountll(m));
                slot = __builtin_amdgcn_readlane(slot, lead);
            }
            qwave += __builtin_popcountll(m);
            if (pend) {
                const int k = slot + static_cast<int>(__builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u)));
                qd[k] = d;
                qk[k] = static_cast<uint16_t>((tid << 3) | yi);
            }
        }
        if (!pend) {
            const int state = s < 0 ? mctgen::kDefaultBlock : s;
            out[yi * 256] = static_cast<uint8_t>(state);
            if (!((mctgen::kAirMask >> state) & 1u)) top = yi;
            frag |= static_cast<uint32_t>(block_class(state)) << (2 * yi);
        }
        for (int k = 0; k < mctgen::kNumChannels; ++k) val[k] += step[k];
    }
    hm[tid] = top;
    fr[tid] = frag;   // queued blocks still at 0, kClsSolid
    if (one_wave && tid == 0) *qn = qwave;
    __syncthreads();
    const int q = *qn;
    for (int e = tid; e < q; e += W) {
        const int key = qk[e], lt = key >> 3, yi = key & 7;
        const int ex = (col0 + lt) & 15, ez = (col0 + lt) >> 4;
        const int ebx = cr.x * 16 + ex, eby = ylo + yi, ebz = cr.z * 16 + ez;
        const AqNearest nr = aquifer_nearest(C, cr.x, cr.z, ebx, eby, ebz);
        const int s = aquifer_back(O, C, ebx, eby, ebz, static_cast<double>(qd[e]), nr);
        const int state = s < 0 ? mctgen::kDefaultBlock : s;
        A.blocks[static_cast<size_t>(c) * kPer + static_cast<size_t>((cy * 8 + yi) * 256 + ez * 16 + ex)] = static_cast<uint8_t>(state);
        if (!((mctgen::kAirMask >> state) & 1u)) atomicMax(&hm[lt], yi);
        const uint32_t cl = static_cast<uint32_t>(block_class(state));
        if (cl != kClsSolid) atomicOr(&fr[lt], cl << (2 * yi));
    }
    __syncthreads();
    top = hm[tid];
    if (A.clsw)
        reinterpret_cast<uint16_t*>(A.clsw + (static_cast<size_t>(c) * (kHeight / 16) + (cy >> 1)) * 256 + col0 + tid)[cy & 1] = static_cast<uint16_t>(fr[tid]);
    // WORLD_SURFACE_WG after the fill: the highest non-air block + 1 of the column, the max over
    // its cells (the heights start at minY; a plain read skips most of the atomics)
    if (top >= 0 && A.h0) {
        int* h = A.h0 + static_cast<size_t>(c) * 256 + x * 16 + z;
        const int v = ylo + top + 1;
        if (v > *reinterpret_cast<volatile int*>(h)) atomicMax(h, v);
    }
}

// Heights start at minY before the fill raises them; the exact-pass chunk list starts empty.
__global__ __launch_bounds__(256) void k_heights_init(int n, int* h0, int* redo_count) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx == 0) *redo_count = 0;
    if (idx < n * 256) h0[idx] = kMinY;
}

__global__ __launch_bounds__(256) void k_fill_core(FillArgs A, int n) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * 256 * kCellsY) return;
    fill_cell_core(A, idx / (256 * kCellsY), idx % (256 * kCellsY));
}

// Column records of the surface pass (2D rule noises, snow-line noise, frozen-patch flag).
__global__ __launch_bounds__(256) void k_surface_cols(const SeedDev* seeds, const ChunkRef* chunks, int n, float* rec, const int* nb, const uint64_t* bmask) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * 256) return;
    surface_column(seeds, c_model, chunks, rec, idx, nb, bmask);
}

// Ore-vein outcomes of the blocks the ore rules can reach, ahead of the surface walk: one thread
// per block column, adjacent threads adjacent x.
__global__ __launch_bounds__(256) void k_ore(SurfArgs A, int n) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * 256) return;
    A.M = &c_model;
    ore_precompute(A, idx);
}

// Threads of a wave take adjacent x so the column walks read adjacent bytes of a y-layer.
__global__ __launch_bounds__(64) void k_surface1(SurfArgs A, int n, const int* h0, int* h1, int* h2) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * 256) return;
    A.M = &c_model;
    const int c = idx >> 8, t = idx & 255;
    surface_pass1(A, h0, h1, h2, c, (t & 15) * 16 + (t >> 4));
}

// A workgroup is 64 columns of one chunk (x 0..15 by 4 z rows, lane = (z & 3) * 16 + x). What
// the walk reads of the blocks is their class (solid / air / water / other fluid): each lane
// packs its column's classes into 24 LDS words (2 bits per block, word w of lane l at
// w * 64 + l, so a y step of the wave reads 64 consecutive words), 6 KB per workgroup instead
// of the 24 KB of the column bytes; rule results go straight to the chunk in global memory.
static_assert(kHeight % 16 == 0, "class words hold 16 blocks");
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_surface2(SurfArgs A, int n, const int* h0, const int* h1, const int* h2, int* hf) {
    __shared__ uint32_t cls[(kHeight / 16) * 64];
    __shared__ float recs[kColF * 64];
    // the chunk index comes from the workgroup index alone, so everything derived from the chunk
    // (seed tables, column records, neighbour biome chunks) is wave-uniform
    const int c = static_cast<int>(blockIdx.x >> 2);
    if (c >= n) return;   // whole workgroups (256 columns per chunk)
    A.M = &c_model;
    const int t = static_cast<int>((blockIdx.x & 3) * 64 + threadIdx.x);
    const int x = t & 15, z = t >> 4;
    const uint32_t* cw = A.clsw + static_cast<size_t>(c) * (kHeight / 16) * 256 + t;
    for (int w = 0; w < kHeight / 16; ++w) cls[w * 64 + threadIdx.x] = cw[w * 256];
    const float* rec = A.colrec + (static_cast<size_t>(c) * 256 + static_cast<size_t>(x * 16 + z)) * kColF;
    for (int k = 0; k < kColF; ++k) recs[k * 64 + threadIdx.x] = rec[k];
    surface_pass2<true>(A, h0, h1, h2, hf, c, x * 16 + z, (uint32_t MCT_LDS*)(cls + threadIdx.x), (const float MCT_LDS*)(recs + threadIdx.x));
}

// A chunk needs the exact pass if any column's final height differs from the height the
// speculative pass assumed for it.
__global__ __launch_bounds__(256) void k_surface_check(int n, const int* h2, const int* hf, int* redo, int* redo_count, unsigned long long* stats) {
    const int c = blockIdx.x;
    if (c >= n) return;
    const int t = threadIdx.x;
    const bool bad = h2[c * 256 + t] != hf[c * 256 + t];
    const int any = __syncthreads_or(bad ? 1 : 0);
    if (t == 0 && any) {
        redo[atomicAdd(redo_count, 1)] = c;
        atomicAdd(stats + 3, 1ull);
    }
}

// Exact pass for flagged chunks: refill the chunk, recompute its heights, then process its
// columns by anti-diagonals (x + z = k): a column reads its x-1 and z-1 neighbours after their
// full processing and its x+1 and z+1 neighbours before theirs, as the x-major loop does.
__global__ __launch_bounds__(256) void k_surface_exact(FillArgs F, SurfArgs A, const int* redo, const int* redo_count) {
    A.M = &c_model;
    __shared__ int hm[256];
    const int total = *redo_count;
    for (int r = blockIdx.x; r < total; r += gridDim.x) {
    const int c = redo[r];
    __syncthreads();
    for (int t = threadIdx.x; t < 256 * kCellsY; t += blockDim.x) fill_cell(F, c, t, nullptr);
    __syncthreads();
    {
        const int t = threadIdx.x;
        hm[t] = column_height(c_model, F.blocks + static_cast<size_t>(c) * kPer, t >> 4, t & 15);
    }
    __syncthreads();
    for (int k = 0; k <= 30; ++k) {
        const int t = threadIdx.x;
        const int lx = t >> 4, lz = t & 15;
        if (lx + lz == k) surface_exact_column(A, hm, c, t);
        __syncthreads();
    }
    }
}

__global__ __launch_bounds__(256) void k_gather_biomes(int n, const int* nb, const uint8_t* bb, uint8_t* out) {
    const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<size_t>(n) * kQPer) return;
    const size_t c = idx / kQPer, q = idx % kQPer;
    out[idx] = bb[static_cast<size_t>(nb[c * 9 + 4]) * kQPer + q];
}

// Biome-only queries: one block per thread (seed slot, x, y, z); 255 where the float search
// cannot serve it (counted in *unserved).
__global__ __launch_bounds__(256) void k_biome_query(const SeedDev* seeds, const int32_t* slot, const int32_t* xyz, int n, uint8_t* out, int* unserved) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int b = biome_query(seeds[slot[i]], c_model, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    if (b < 0) atomicAdd(unserved, 1);
    out[i] = static_cast<uint8_t>(b < 0 ? 255 : b);
}

// The column-cached zoom of the surface pass against zoom_pick: one pseudo-random column per
// thread (x, z uniform in +-29,999,984), every y. out: checked, mismatches, row refills.
__global__ __launch_bounds__(256) void k_zoom_selftest(int64_t zoom_seed, uint64_t salt, int columns, unsigned long long* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= columns) return;
    uint64_t h = salt + 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(i + 1);
    h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
    h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
    h ^= h >> 31;
    const int x = static_cast<int>((h & 0xFFFFFFFFull) % 59999969ull) - 29999984;
    const int z = static_cast<int>((h >> 32) % 59999969ull) - 29999984;
    ZoomColumn zc;
    zc.init(zoom_seed, x, z);
    unsigned long long bad = 0;
    for (int y = kMinY + kHeight - 1; y >= kMinY; --y) {
        int a, b, c, qa, qb, qc;
        zc.pick(y, a, b, c);
        zoom_pick(zoom_seed, x, y, z, qa, qb, qc);
        bad += (a != qa || b != qb || c != qc) ? 1u : 0u;
    }
    atomicAdd(out, static_cast<unsigned long long>(kHeight));
    if (bad) atomicAdd(out + 1, bad);
    atomicAdd(out + 2, static_cast<unsigned long long>(zc.refills));
}

// Top-down record of each block column of n chunks (one thread per column, t = z * 16 + x):
// the highest non-air block below the top, then under fluid the first non-fluid block.
// 8 bytes: top state, floor state, min(255, top_y - floor_y), 0, int16 top y, int16 floor y
// (little-endian), the record of the reference bench's MCTD view (bench/bench_main.cpp:132-166).
__global__ __launch_bounds__(256) void k_topdown(const uint8_t* blocks, int n, uint8_t* out) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * 256) return;
    const uint8_t* b = blocks + static_cast<size_t>(idx >> 8) * kPer + (idx & 255);
    int y = kHeight - 1;
    while (y > 0 && ((mctgen::kAirMask >> b[y * 256]) & 1u)) --y;
    const int top = b[y * 256], top_y = y;
    while (y > 0 && ((mctgen::kFluidMask >> b[y * 256]) & 1u)) --y;
    const int floor_s = b[y * 256];
    const int ty = kMinY + top_y, fy = kMinY + y;
    const uint32_t lo = static_cast<uint32_t>(top) | (static_cast<uint32_t>(floor_s) << 8) | (static_cast<uint32_t>(min(255, top_y - y)) << 16);
    const uint32_t hi = (static_cast<uint32_t>(ty) & 0xFFFFu) | ((static_cast<uint32_t>(fy) & 0xFFFFu) << 16);
    reinterpret_cast<uint2*>(out)[idx] = make_uint2(lo, hi);
}

// Per-chunk 64-bit digest of output bytes in HBM, one workgroup per chunk: the sum (mod 2^64)
// over 8-byte words w_i of mix(w_i + (i + 1) * golden), finalised with the byte count. Used to
// compare whole runs without copying them to the host.
__device__ inline uint64_t digest_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_chunk_digest(const uint64_t* data, int64_t words_per_chunk, uint64_t* out) {
    const uint64_t* p = data + static_cast<size_t>(blockIdx.x) * static_cast<size_t>(words_per_chunk);
    uint64_t h = 0;
    for (int64_t i = threadIdx.x; i < words_per_chunk; i += 256) h += digest_mix(p[i] + 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(i + 1));
    __shared__ uint64_t part[256];
    part[threadIdx.x] = h;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = digest_mix(part[0] ^ static_cast<uint64_t>(words_per_chunk * 8));
}

__global__ __launch_bounds__(256) void k_copy16(size_t n16, const uint4* src, uint4* dst) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n16) dst[i] = src[i];
}

// ---- dedupe: LSD radix sort of (key, index), run heads, scan, scatter ------------------------
constexpr int kRsThreads = 256, kRsItems = 16, kRsTile = kRsThreads * kRsItems;
constexpr int kRsWaves = kRsThreads / 64, kRsRounds = kRsItems;   // per wave: kRsRounds rounds of 64 keys

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint64_t* keys, int n, int shift, int nblocks, int* hist) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRsTile;
    for (int k = 0; k < kRsItems; ++k) {
        const int i = base + k * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of `total` ints in one workgroup (each thread a contiguous run).
__global__ __launch_bounds__(1024) void k_scan_small(int* v, int total, int* total_out) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (total + 1023) / 1024;
    const int b = t * per, e = b + per < total ? b + per : total;
    int s = 0;
    for (int i = b; i < e; ++i) s += v[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int a = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += a;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = b; i < e; ++i) {
        const int x = v[i];
        v[i] = run;
        run += x;
    }
    if (t == 1023 && total_out) *total_out = part[1023];
}

// Exclusive scan of long arrays (the radix histograms) in three launches: per-tile sums, a scan
// of the tile sums, then each tile scanned in LDS from its offset. Integer sums: the same
// result as any other order.
constexpr int kScanTile = 2048;   // 256 threads x 8
__global__ __launch_bounds__(256) void k_scan_tiles(const int* v, int total, int* sums) {
    __shared__ int part[256];
    const int base = blockIdx.x * kScanTile;
    int a = 0;
    for (int j = 0; j < 8; ++j) {
        const int i = base + j * 256 + static_cast<int>(threadIdx.x);
        if (i < total) a += v[i];
    }
    part[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = part[0];
}

__global__ __launch_bounds__(256) void k_scan_apply(int* v, int total, const int* offsets) {
    __shared__ int tile[kScanTile];
    __shared__ int ts[256];
    const int base = blockIdx.x * kScanTile, t = static_cast<int>(threadIdx.x);
    for (int j = 0; j < 8; ++j) {
        const int i = base + j * 256 + t;
        tile[j * 256 + t] = i < total ? v[i] : 0;
    }
    __syncthreads();
    int local = 0;
    for (int k = 0; k < 8; ++k) local += tile[t * 8 + k];
    ts[t] = local;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int a = t >= off ? ts[t - off] : 0;
        __syncthreads();
        ts[t] += a;
        __syncthreads();
    }
    int run = offsets[blockIdx.x] + ts[t] - local;
    for (int k = 0; k < 8; ++k) {
        const int x = tile[t * 8 + k];
        tile[t * 8 + k] = run;
        run += x;
    }
    __syncthreads();
    for (int j = 0; j < 8; ++j) {
        const int i = base + j * 256 + t;
        if (i < total) v[i] = tile[j * 256 + t];
    }
}

__device__ __forceinline__ unsigned long long match_digit(int d, bool valid) {
    unsigned long long m = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
        const unsigned long long bb = __ballot(valid && ((d >> b) & 1));
        m &= ((d >> b) & 1) ? bb : ~bb;
    }
    return m;
}

// Stable scatter: wave w owns keys [base + w*512, +512) in rounds of 64; ranks come from
// 64-lane ballots, per-wave digit counts from a first pass.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint64_t* kin, const int* vin, int n, int shift, int nblocks, const int* hist,
                                                          uint64_t* kout, int* vout) {
    __shared__ int cnt[kRsWaves][256];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < kRsWaves * 256; i += kRsThreads) (&cnt[0][0])[i] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRsTile + w * (kRsRounds * 64);
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = 0; r < kRsRounds; ++r) {
        const int i = base + r * 64 + lane;
        const bool valid = i < n;
        const int d = valid ? static_cast<int>((kin[i] >> shift) & 255) : 0;
        const unsigned long long peers = match_digit(d, valid);
        if (valid && (peers & below) == 0) cnt[w][d] += __popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    {
        const int d = threadIdx.x;   // 256 threads, one digit each
        int run = hist[d * nblocks + blockIdx.x];
        for (int ww = 0; ww < kRsWaves; ++ww) {
            const int c = cnt[ww][d];
            cnt[ww][d] = run;
            run += c;
        }
    }
    __syncthreads();
    for (int r = 0; r < kRsRounds; ++r) {
        const int i = base + r * 64 + lane;
        const bool valid = i < n;
        const uint64_t key = valid ? kin[i] : 0;
        const int d = static_cast<int>((key >> shift) & 255);
        const unsigned long long peers = match_digit(d, valid);
        if (valid) {
            const int pos = cnt[w][d] + __popcll(peers & below);
            kout[pos] = key;
            vout[pos] = vin ? vin[i] : i;
        }
        __builtin_amdgcn_wave_barrier();
        if (valid && (peers & below) == 0) cnt[w][d] += __popcll(peers);
        __builtin_amdgcn_wave_barrier();
    }
}

__global__ __launch_bounds__(256) void k_heads(const uint64_t* sorted, int n, int* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = (i == 0 || sorted[i] != sorted[i - 1]) ? 1 : 0;
}

// Per-tile sums of the head flags (tile = 2048), then in-tile prefix + tile offset.
__global__ __launch_bounds__(256) void k_tile_sums(const int* flag, int n, int* sums) {
    __shared__ int s[256];
    const int base = blockIdx.x * kRsTile;
    int a = 0;
    for (int k = 0; k < kRsItems; ++k) {
        const int i = base + k * 256 + threadIdx.x;
        if (i < n) a += flag[i];
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = s[0];
}

// unique[u] = key at each run head, map[index] = u, with u from the exclusive tile offsets.
__global__ __launch_bounds__(256) void k_unique_scatter(const uint64_t* sorted, const int* idx, const int* flag, int n, const int* tile_off,
                                                        uint64_t* unique, int* map, int* count, const int* total) {
    __shared__ int s[kRsTile];
    const int base = blockIdx.x * kRsTile;
    for (int k = 0; k < kRsItems; ++k) {
        const int i = base + k * 256 + threadIdx.x;
        s[k * 256 + threadIdx.x] = i < n ? flag[i] : 0;
    }
    __syncthreads();
    // inclusive scan of the tile's 2048 flags: each thread scans 8 consecutive, then a
    // 256-wide scan of the thread sums
    __shared__ int ts[256];
    int local[kRsItems];
    int a = 0;
    for (int k = 0; k < kRsItems; ++k) {
        a += s[threadIdx.x * kRsItems + k];
        local[k] = a;
    }
    ts[threadIdx.x] = a;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = threadIdx.x >= off ? ts[threadIdx.x - off] : 0;
        __syncthreads();
        ts[threadIdx.x] += v;
        __syncthreads();
    }
    const int before = ts[threadIdx.x] - a + tile_off[blockIdx.x];
    for (int k = 0; k < kRsItems; ++k) {
        const int i = base + threadIdx.x * kRsItems + k;
        if (i < n) {
            const int u = before + local[k] - 1;
            if (s[threadIdx.x * kRsItems + k]) unique[u] = sorted[i];
            map[idx[i]] = u;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *count = *total;
}

KeyCodec codec_for(const ChunkRef* ch, int n, int xlo, int xhi, int zlo, int zhi) {
    int mnx = INT32_MAX, mxx = INT32_MIN, mnz = INT32_MAX, mxz = INT32_MIN, mns = INT32_MAX, mxs = INT32_MIN;
    for (int i = 0; i < n; ++i) {
        mnx = std::min(mnx, ch[i].x); mxx = std::max(mxx, ch[i].x); mnz = std::min(mnz, ch[i].z); mxz = std::max(mxz, ch[i].z);
        mns = std::min(mns, ch[i].seed); mxs = std::max(mxs, ch[i].seed);
    }
    auto bits = [](int64_t range) { int b = 0; while ((int64_t{1} << b) < range) ++b; return b; };
    KeyCodec k{};
    k.x0 = 4 * mnx + xlo;
    k.z0 = 4 * mnz + zlo;
    k.s0 = mns;
    k.bx = bits(static_cast<int64_t>(4 * mxx + xhi) - k.x0 + 1);
    k.bz = bits(static_cast<int64_t>(4 * mxz + zhi) - k.z0 + 1);
    k.bits = bits(static_cast<int64_t>(mxs - mns + 1)) + 1 + k.bx + k.bz;
    if (k.bits > 64) throw Error(MCT_ECAPACITY, "chunk range too wide for 64-bit position keys");
    return k;
}

unsigned grid_of(size_t threads, int block) { return static_cast<unsigned>((threads + block - 1) / block); }

// Device dedupe of 64-bit keys: unique[0..*count) ascending, map[i] = index of keys[i].
struct Dedupe {
    int cap = 0;
    uint64_t *ka = nullptr, *kb = nullptr;
    int *va = nullptr, *vb = nullptr, *hist = nullptr, *flag = nullptr, *tsum = nullptr, *total = nullptr, *hsum = nullptr;
    explicit Dedupe(int capacity) : cap(capacity) {
        const int nb = (cap + kRsTile - 1) / kRsTile;
        hip_check(hipMalloc(&hsum, sizeof(int) * ((256 * nb + kScanTile - 1) / kScanTile + 1)), "malloc");
        hip_check(hipMalloc(&ka, sizeof(uint64_t) * cap), "malloc");
        hip_check(hipMalloc(&kb, sizeof(uint64_t) * cap), "malloc");
        hip_check(hipMalloc(&va, sizeof(int) * cap), "malloc");
        hip_check(hipMalloc(&vb, sizeof(int) * cap), "malloc");
        hip_check(hipMalloc(&flag, sizeof(int) * cap), "malloc");
        hip_check(hipMalloc(&hist, sizeof(int) * 256 * nb), "malloc");
        hip_check(hipMalloc(&tsum, sizeof(int) * (nb + 1)), "malloc");
        hip_check(hipMalloc(&total, sizeof(int)), "malloc");
    }
    ~Dedupe() {
        void* p[] = {ka, kb, va, vb, flag, hist, tsum, total, hsum};
        for (void* q : p) (void)hipFree(q);
    }
    void run(const uint64_t* keys, int n, int key_bits, uint64_t* unique, int* map, int* count, hipStream_t s) {
        if (n > cap) throw Error(MCT_ECAPACITY, "dedupe capacity exceeded");
        if (n == 0) return;
        const int nb = (n + kRsTile - 1) / kRsTile;
        const uint64_t* kin = keys;
        const int* vin = nullptr;
        uint64_t* kout = ka;
        int* vout = va;
        for (int shift = 0; shift < key_bits; shift += 8) {
            k_rs_hist<<<nb, kRsThreads, 0, s>>>(kin, n, shift, nb, hist);
            const int tiles = (256 * nb + kScanTile - 1) / kScanTile;
            if (tiles == 1) {
                k_scan_small<<<1, 1024, 0, s>>>(hist, 256 * nb, nullptr);
            } else {
                k_scan_tiles<<<tiles, 256, 0, s>>>(hist, 256 * nb, hsum);
                k_scan_small<<<1, 1024, 0, s>>>(hsum, tiles, nullptr);
                k_scan_apply<<<tiles, 256, 0, s>>>(hist, 256 * nb, hsum);
            }
            k_rs_scatter<<<nb, kRsThreads, 0, s>>>(kin, vin, n, shift, nb, hist, kout, vout);
            kin = kout;
            vin = vout;
            kout = kout == ka ? kb : ka;
            vout = vout == va ? vb : va;
        }
        k_heads<<<grid_of(n, 256), 256, 0, s>>>(kin, n, flag);
        k_tile_sums<<<nb, 256, 0, s>>>(flag, n, tsum);
        k_scan_small<<<1, 1024, 0, s>>>(tsum, nb, total);
        k_unique_scatter<<<nb, 256, 0, s>>>(kin, vin, flag, n, tsum, unique, map, count, total);
        hip_check(hipGetLastError(), "dedupe kernels");
    }
};

// ---- generator -------------------------------------------------------------------------------
// Kernel groups timed with HIP events on the batch stream when profiling is enabled.
enum KernelId {
    kKBiomeColumns, kKBiomeQuarts, kKBiomeTies, kKBiomeChain, kKDedupeCols, kKColumns, kKCorners, kKDedupeGrid, kKFts, kKGridScatter,
    kKAqCells, kKFill, kKHeights, kKSurface1, kKSurface2, kKSurfaceExact, kKGather, kKSurfaceCols, kKOre, kKCount
};
const char* const kKernelNames[kKCount] = {"biome_columns", "biome_quarts", "biome_ties", "biome_chain", "dedupe_columns", "columns", "corners",
                                           "dedupe_grid", "fts", "grid_scatter+aquifer_skip", "aquifer_cells", "fill", "heights", "surface1",
                                           "surface2", "surface_check+exact", "gather_biomes", "surface_columns", "ore_veins"};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev[9] = {};
    bool pending = false;
    int batch = -1;
    int cap;                        // chunks
    int bcap = 0;                   // biome chunks
    size_t bytes = 0;               // device memory held (without the biome-chunk buffers)
    // pinned staging
    ChunkRef *h_ch = nullptr, *h_bch = nullptr;
    int* h_nb = nullptr;
    // device
    ChunkRef *d_ch = nullptr, *d_bch = nullptr;
    int* d_nb = nullptr;
    float *d_bcol = nullptr, *d_colv = nullptr, *d_corners = nullptr, *d_gv = nullptr, *d_prelim = nullptr;
    int64_t *d_targets = nullptr, *d_mind = nullptr;
    int32_t *d_lstar = nullptr, *d_grid = nullptr, *d_cloc = nullptr, *d_cfl = nullptr;
    uint8_t *d_tie = nullptr, *d_bb = nullptr, *d_blocks = nullptr, *d_bout = nullptr, *d_aq = nullptr;
    uint64_t* d_bmask = nullptr;    // biome chunks x 16 quart columns x 2: the biomes of each quart column
    GroupRes* d_gres = nullptr;     // biome chunks x 16 quart columns x groups: column search results
    uint64_t *d_ckeys = nullptr, *d_ucols = nullptr, *d_gkeys = nullptr, *d_ug = nullptr;
    int *d_colmap = nullptr, *d_gmap = nullptr, *d_counts = nullptr, *d_skip = nullptr;
    int *d_h0 = nullptr, *d_h1 = nullptr, *d_h2 = nullptr, *d_hf = nullptr, *d_flags = nullptr;
    float* d_colrec = nullptr;
    uint8_t* d_ore = nullptr;
    uint32_t* d_clsw = nullptr;     // block-class words (SurfArgs::clsw)
    float* d_density = nullptr;     // core configuration's density output (allocated on first use)
    std::unique_ptr<Dedupe> dedupe;
    hipEvent_t kev[kKCount][2] = {};
    bool kev_recorded = false;
    // destination of the batch being generated
    int n = 0;
    uint8_t *out_terrain = nullptr, *out_biomes = nullptr, *out_aquifer = nullptr;
    bool to_device = false;

    explicit Slot(int mb) : cap(mb) {
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
        for (auto& e : ev) hip_check(hipEventCreate(&e), "event");
        hip_check(hipHostMalloc(&h_ch, sizeof(ChunkRef) * mb, hipHostMallocDefault), "pinned");
        hip_check(hipHostMalloc(&h_bch, sizeof(ChunkRef) * mb * 9, hipHostMallocDefault), "pinned");
        hip_check(hipHostMalloc(&h_nb, sizeof(int) * mb * 9, hipHostMallocDefault), "pinned");
        auto M = [this](auto** p, size_t bytes) {
            hip_check(hipMalloc(reinterpret_cast<void**>(p), bytes > 0 ? bytes : 16), "malloc");
            this->bytes += bytes;
        };
        const size_t ncol = std::max(mctgen::kNumColumns, 1);
        M(&d_ch, sizeof(ChunkRef) * mb);
        M(&d_nb, sizeof(int) * mb * 9);
        M(&d_ckeys, sizeof(uint64_t) * mb * kCols);
        M(&d_ucols, sizeof(uint64_t) * mb * kCols);
        M(&d_colmap, sizeof(int) * mb * kCols);
        M(&d_colv, sizeof(float) * mb * kCols * ncol);
        M(&d_corners, sizeof(float) * mb * kCols * kCh * kRows);
        M(&d_gkeys, sizeof(uint64_t) * mb * kGridKeys);
        M(&d_ug, sizeof(uint64_t) * mb * kGridKeys);
        M(&d_gmap, sizeof(int) * mb * kGridKeys);
        M(&d_gv, sizeof(float) * mb * kGridKeys);
        M(&d_grid, sizeof(int32_t) * mb * kGridCells);
        M(&d_prelim, sizeof(float) * mb * 4);
        M(&d_skip, sizeof(int) * mb);
        M(&d_cloc, sizeof(int32_t) * mb * kAqCells * 4);
        M(&d_cfl, sizeof(int32_t) * mb * kAqCells * 2);
        M(&d_blocks, static_cast<size_t>(mb) * kPer);
        M(&d_bout, static_cast<size_t>(mb) * kQPer);
        M(&d_h0, sizeof(int) * mb * 256);
        M(&d_h1, sizeof(int) * mb * 256);
        M(&d_h2, sizeof(int) * mb * 256);
        M(&d_hf, sizeof(int) * mb * 256);
        M(&d_colrec, sizeof(float) * mb * 256 * kColF);
        M(&d_ore, static_cast<size_t>(mb) * 256 * kOreRowsPad);
        M(&d_clsw, sizeof(uint32_t) * mb * (kHeight / 16) * 256);
        M(&d_flags, sizeof(int) * mb);
        M(&d_counts, sizeof(int) * 4);
        dedupe = std::make_unique<Dedupe>(mb * kGridKeys);
    }
    ~Slot() {
        if (stream) (void)hipStreamSynchronize(stream);
        void* p[] = {d_ch, d_bch, d_nb, d_bcol, d_colv, d_corners, d_gv, d_prelim, d_targets, d_mind, d_lstar, d_grid, d_cloc, d_cfl, d_tie, d_bb, d_bmask, d_gres,
                     d_blocks, d_bout, d_aq, d_ckeys, d_ucols, d_gkeys, d_ug, d_colmap, d_gmap, d_counts, d_skip, d_h0, d_h1, d_h2, d_hf, d_flags, d_colrec,
                     d_ore, d_clsw, d_density};
        for (void* q : p) (void)hipFree(q);
        dedupe.reset();
        (void)hipHostFree(h_ch);
        (void)hipHostFree(h_bch);
        (void)hipHostFree(h_nb);
        for (auto& e : ev) (void)hipEventDestroy(e);
        for (auto& k : kev)
            for (auto& e : k)
                if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void ensure_biome_capacity(int nbk) {
        if (nbk <= bcap) return;
        void* old[] = {d_bch, d_bcol, d_targets, d_mind, d_lstar, d_tie, d_bb, d_bmask, d_gres};
        for (void* q : old) (void)hipFree(q);
        const size_t q = static_cast<size_t>(nbk) * kQPer;
        hip_check(hipMalloc(&d_bch, sizeof(ChunkRef) * nbk), "malloc");
        hip_check(hipMalloc(&d_bcol, sizeof(float) * nbk * 16 * std::max(mctgen::kNumClimateColumns, 1)), "malloc");
        hip_check(hipMalloc(&d_targets, sizeof(int64_t) * q * 7), "malloc");
        hip_check(hipMalloc(&d_mind, sizeof(int64_t) * q), "malloc");
        hip_check(hipMalloc(&d_lstar, sizeof(int32_t) * q), "malloc");
        hip_check(hipMalloc(&d_tie, q), "malloc");
        hip_check(hipMalloc(&d_bb, q), "malloc");
        hip_check(hipMalloc(&d_bmask, sizeof(uint64_t) * nbk * 16 * 2), "malloc");
        hip_check(hipMalloc(&d_gres, sizeof(GroupRes) * nbk * 16 * kColGroups), "malloc");
        bcap = nbk;
    }
    size_t biome_bytes() const {
        const size_t q = static_cast<size_t>(bcap) * kQPer;
        return sizeof(ChunkRef) * bcap + sizeof(float) * bcap * 16 * std::max(mctgen::kNumClimateColumns, 1) + q * (7 * 8 + 8 + 4 + 1 + 1) +
               sizeof(uint64_t) * bcap * 16 * 2 + sizeof(GroupRes) * bcap * 16 * kColGroups;
    }
};

struct Generator {
    int max_batch, nslots, seed_cap;
    std::vector<std::unique_ptr<Slot>> slots;
    size_t next_slot = 0;
    // model
    TreeNodeDev* d_tree = nullptr;
    TreeNodeF* d_treef = nullptr;
    uint8_t* d_leaf_biome = nullptr;
    int32_t* d_leaf_node = nullptr;
    int32_t* d_leaf_rank = nullptr;
    int vary_groups = 0;            // column search groups of the one varying climate coordinate (0: none)
    int* d_blended_first = nullptr;
    int nnormal = 0;
    // resident seeds
    SeedDev* d_seeds = nullptr;
    std::vector<int64_t> slot_seed;
    std::vector<uint64_t> slot_used;
    std::unordered_map<int64_t, int> resident;
    uint64_t use_clock = 0;
    SeedJob* d_jobs = nullptr;
    SeedJob* h_jobs = nullptr;
    int64_t current_seed = 0;
    unsigned long long* d_stats = nullptr;   // [0] quarts, [1] overrides, [2] biome changes, [3] exact redo
    uint64_t quarts = 0;
    hipEvent_t t0 = nullptr;
    std::vector<double> done_ms;
    bool profile = false;
    double kernel_ms[kKCount] = {};
    // Launch shapes (mct_autotune / mct_set_tuning): workgroup size per kernel group (0: fixed
    // by the kernel's layout) and quarts per R-tree search thread. Results do not depend on them.
    int wg[kKCount] = {};
    int run_len = kBiomeRunDefault;
    int W(int k) const { return wg[k] > 0 ? wg[k] : 256; }
    int64_t kernel_launches[kKCount] = {};

    // Records kernel-group events around `launch` on the slot stream when profiling.
    template <typename F>
    void timed(Slot& S, int id, F&& launch) {
        if (!profile) { launch(); return; }
        if (!S.kev[id][0]) {
            hip_check(hipEventCreate(&S.kev[id][0]), "event");
            hip_check(hipEventCreate(&S.kev[id][1]), "event");
        }
        hip_check(hipEventRecord(S.kev[id][0], S.stream), "event");
        launch();
        hip_check(hipEventRecord(S.kev[id][1], S.stream), "event");
        S.kev_recorded = true;
    }

    Generator(int mb, int ns, int sc) : max_batch(mb), nslots(ns), seed_cap(sc) {
        std::vector<TreeNodeDev> tree(mctgen::kTreeNodes);
        for (int i = 0; i < mctgen::kTreeNodes; ++i) {
            const mctgen::TreeNode& t = mctgen::kTree[i];
            for (int k = 0; k < 7; ++k) { tree[i].lo[k] = t.lo[k]; tree[i].hi[k] = t.hi[k]; }
            tree[i].first = t.first;
            tree[i].count = t.count;
            tree[i].leaf = t.leaf;
            tree[i].pad = 0;
        }
        hip_check(hipMalloc(&d_tree, sizeof(TreeNodeDev) * tree.size()), "malloc");
        hip_check(hipMemcpy(d_tree, tree.data(), sizeof(TreeNodeDev) * tree.size(), hipMemcpyHostToDevice), "upload tree");
        const std::vector<TreeNodeF> treef = float_tree(tree);
        if (!treef.empty() && tree_levels(tree) <= kSearchDepth) {
            hip_check(hipMalloc(&d_treef, sizeof(TreeNodeF) * treef.size()), "malloc");
            hip_check(hipMemcpy(d_treef, treef.data(), sizeof(TreeNodeF) * treef.size(), hipMemcpyHostToDevice), "upload tree");
        }
        hip_check(hipMalloc(&d_leaf_biome, mctgen::kTreeLeaves), "malloc");
        hip_check(hipMemcpy(d_leaf_biome, mctgen::kLeafBiome, mctgen::kTreeLeaves, hipMemcpyHostToDevice), "upload");
        hip_check(hipMalloc(&d_leaf_node, sizeof(int32_t) * mctgen::kTreeLeaves), "malloc");
        hip_check(hipMemcpy(d_leaf_node, mctgen::kLeafNode, sizeof(int32_t) * mctgen::kTreeLeaves, hipMemcpyHostToDevice), "upload");
        const std::vector<int32_t> rank = leaf_dfs_rank(tree, mctgen::kTreeLeaves);
        hip_check(hipMalloc(&d_leaf_rank, sizeof(int32_t) * rank.size()), "malloc");
        hip_check(hipMemcpy(d_leaf_rank, rank.data(), sizeof(int32_t) * rank.size(), hipMemcpyHostToDevice), "upload");
        ModelDev M{};
        M.tree = d_tree;
        M.treef = d_treef;
        M.leaf_biome = d_leaf_biome;
        M.leaf_node = d_leaf_node;
        M.leaf_rank = d_leaf_rank;
        if (d_treef) column_groups(treef, M);
        vary_groups = (d_treef && kOneVarying) ? M.col_groups[kVaryCoord] : 0;
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
        hip_check(hipMalloc(&d_blended_first, sizeof(int) * bf.size()), "malloc");
        hip_check(hipMemcpy(d_blended_first, bf.data(), sizeof(int) * bf.size(), hipMemcpyHostToDevice), "upload");
        hip_check(hipMalloc(&d_seeds, sizeof(SeedDev) * seed_cap), "malloc seed tables");
        slot_seed.assign(seed_cap, 0);
        slot_used.assign(seed_cap, 0);
        hip_check(hipMalloc(&d_jobs, sizeof(SeedJob) * seed_cap), "malloc");
        hip_check(hipHostMalloc(&h_jobs, sizeof(SeedJob) * seed_cap, hipHostMallocDefault), "pinned");
        hip_check(hipMalloc(&d_stats, sizeof(unsigned long long) * 4), "malloc");
        hip_check(hipMemset(d_stats, 0, sizeof(unsigned long long) * 4), "memset");
        hip_check(hipEventCreate(&t0), "event");
        for (int i = 0; i < nslots; ++i) slots.push_back(std::make_unique<Slot>(mb));
        hip_check(hipDeviceSynchronize(), "init");
    }

    // Device memory held: seed tables, model tables and every slot (biome buffers at their
    // current capacity).
    size_t device_bytes() const {
        size_t b = sizeof(SeedDev) * seed_cap + sizeof(TreeNodeDev) * mctgen::kTreeNodes + sizeof(TreeNodeF) * mctgen::kTreeNodes;
        for (const auto& s : slots) b += s->bytes + s->biome_bytes();
        return b;
    }

    ~Generator() {
        (void)hipDeviceSynchronize();
        slots.clear();
        void* q[] = {d_qxyz, d_qslot, d_qbad, d_qout};
        for (void* x : q) (void)hipFree(x);
        void* p[] = {d_tree, d_treef, d_leaf_biome, d_leaf_node, d_leaf_rank, d_blended_first, d_seeds, d_jobs, d_stats};
        for (void* q : p) (void)hipFree(q);
        (void)hipHostFree(h_jobs);
        if (t0) (void)hipEventDestroy(t0);
    }

    void collect(Slot& S, mct_timing* tm) {
        if (!S.pending) return;
        hip_check(hipEventSynchronize(S.ev[8]), "sync");
        S.pending = false;
        if (S.batch >= 0 && S.batch < static_cast<int>(done_ms.size())) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, t0, S.ev[8]), "event time");
            done_ms[static_cast<size_t>(S.batch)] = ms;
        }
        S.batch = -1;
        if (S.kev_recorded) {
            for (int k = 0; k < kKCount; ++k) {
                if (!S.kev[k][0]) continue;
                float ms = 0;
                if (hipEventElapsedTime(&ms, S.kev[k][0], S.kev[k][1]) == hipSuccess) {
                    kernel_ms[k] += ms;
                    kernel_launches[k] += 1;
                }
            }
            S.kev_recorded = false;
        }
        if (!tm) return;
        float t[8];
        for (int i = 0; i < 8; ++i) hip_check(hipEventElapsedTime(&t[i], S.ev[i], S.ev[i + 1]), "event time");
        tm->h2d_ms += t[0];
        tm->biome_ms += t[1];
        tm->noise_ms += t[2];
        tm->aquifer_ms += t[3];
        tm->fill_ms += t[4];
        tm->surface_ms += t[5] + t[6];
        tm->d2h_ms += t[7];
    }
    void collect_all(mct_timing* tm) {
        for (auto& s : slots) collect(*s, tm);
    }

    // Slots of the given seeds (building missing tables on `stream`); at most seed_cap distinct.
    void ensure_seeds(const std::vector<int64_t>& need, hipStream_t stream) {
        int njobs = 0;
        ++use_clock;
        for (int64_t s : need) {
            auto it = resident.find(s);
            if (it != resident.end()) {
                slot_used[it->second] = use_clock;
                continue;
            }
            int victim = -1;
            for (int k = 0; k < seed_cap; ++k)
                if (slot_used[k] != use_clock && (victim < 0 || slot_used[k] < slot_used[victim])) victim = k;
            if (victim < 0) throw Error(MCT_ECAPACITY, "more distinct seeds in one batch than seed_capacity");
            if (slot_used[victim] != 0) {
                collect_all(nullptr);   // in-flight batches may still read the victim's tables
                resident.erase(slot_seed[victim]);
            }
            slot_seed[victim] = s;
            slot_used[victim] = use_clock;
            resident[s] = victim;
            h_jobs[njobs++] = SeedJob{s, victim, 0};
        }
        if (njobs == 0) return;
        hip_check(hipStreamSynchronize(stream), "sync");   // h_jobs staging is shared
        hip_check(hipMemcpyAsync(d_jobs, h_jobs, sizeof(SeedJob) * njobs, hipMemcpyHostToDevice, stream), "upload seed jobs");
        k_seed_layers<<<grid_of(static_cast<size_t>(njobs) * nnormal, kSetupThreads), kSetupThreads, 0, stream>>>(d_seeds, d_jobs, njobs, nnormal);
        k_seed_serial<<<grid_of(static_cast<size_t>(njobs) * (mctgen::kNumBlended + 1), kSetupThreads), kSetupThreads, 0, stream>>>(d_seeds, d_jobs, njobs,
                                                                                                                               d_blended_first);
        hip_check(hipGetLastError(), "seed setup kernels");
        hip_check(hipStreamSynchronize(stream), "seed setup");
    }

    // Core configuration (TerrainGenerator, no biomes / aquifers / material rules): the corner
    // pass, then final density and the core fill per block. fill / density: destinations of this
    // batch (device pointers when to_device, else host), either may be null.
    void run_batch_core(Slot& S, const int64_t* seeds, const int32_t* xz, int n, uint8_t* fill, float* density, bool to_device) {
        S.n = n;
        std::vector<int64_t> distinct;
        {
            std::unordered_map<int64_t, int> seen;
            for (int i = 0; i < n; ++i)
                if (seen.emplace(seeds[i], 1).second) distinct.push_back(seeds[i]);
        }
        ensure_seeds(distinct, S.stream);
        for (int i = 0; i < n; ++i) S.h_ch[i] = ChunkRef{xz[2 * i], xz[2 * i + 1], resident.at(seeds[i]), 0};
        if (density && !to_device && !S.d_density) hip_check(hipMalloc(&S.d_density, sizeof(float) * static_cast<size_t>(S.cap) * kPer), "malloc");
        hipStream_t st = S.stream;
        hip_check(hipEventRecord(S.ev[0], st), "event");
        hip_check(hipMemcpyAsync(S.d_ch, S.h_ch, sizeof(ChunkRef) * n, hipMemcpyHostToDevice, st), "upload chunks");
        hip_check(hipEventRecord(S.ev[1], st), "event");
        hip_check(hipEventRecord(S.ev[2], st), "event");
        const SeedDev* sd = d_seeds;
        const int ck = n * kCols;
        const KeyCodec KC = codec_for(S.h_ch, n, 0, 4, 0, 4);
        k_col_keys<<<grid_of(ck, 256), 256, 0, st>>>(KC, S.d_ch, n, S.d_ckeys);
        timed(S, kKDedupeCols, [&] { S.dedupe->run(S.d_ckeys, ck, KC.bits, S.d_ucols, S.d_colmap, S.d_counts, st); });
        timed(S, kKColumns, [&] { k_columns<<<grid_of(ck, W(kKColumns)), W(kKColumns), 0, st>>>(KC, sd, ck, S.d_ucols, S.d_counts, S.d_colv); });
        timed(S, kKCorners, [&] { k_corners<<<grid_of(ck, kCornerCols), 256, 0, st>>>(KC, sd, ck, S.d_ucols, S.d_counts, S.d_colv, S.d_corners); });
        hip_check(hipGetLastError(), "noise kernels");
        hip_check(hipEventRecord(S.ev[3], st), "event");
        hip_check(hipEventRecord(S.ev[4], st), "event");
        uint8_t* dfill = fill ? (to_device ? fill : S.d_blocks) : nullptr;
        float* ddens = density ? (to_device ? density : S.d_density) : nullptr;
        FillArgs F{sd, S.d_ch, S.d_corners, S.d_colmap, nullptr, nullptr, nullptr, dfill, ddens};
        timed(S, kKFill, [&] { k_fill_core<<<grid_of(static_cast<size_t>(n) * 256 * kCellsY, W(kKFill)), W(kKFill), 0, st>>>(F, n); });
        hip_check(hipGetLastError(), "fill kernel");
        for (int e = 5; e <= 7; ++e) hip_check(hipEventRecord(S.ev[e], st), "event");
        if (!to_device) {
            if (fill) hip_check(hipMemcpyAsync(fill, S.d_blocks, static_cast<size_t>(n) * kPer, hipMemcpyDeviceToHost, st), "download fill");
            if (density) hip_check(hipMemcpyAsync(density, S.d_density, sizeof(float) * static_cast<size_t>(n) * kPer, hipMemcpyDeviceToHost, st), "download density");
        }
        hip_check(hipEventRecord(S.ev[8], st), "event");
        S.pending = true;
    }

    // Enqueues one batch on slot S (uploads, kernels, downloads); returns without waiting.
    void run_batch(Slot& S, const int64_t* seeds, const int32_t* xz, int n, uint8_t* terrain, uint8_t* biomes, uint8_t* aquifer, bool to_device) {
        S.n = n;
        S.out_terrain = terrain;
        S.out_biomes = biomes;
        S.out_aquifer = aquifer;
        S.to_device = to_device;
        if (aquifer && !S.d_aq) hip_check(hipMalloc(&S.d_aq, static_cast<size_t>(S.cap) * kPer), "malloc");
        // seeds of this batch -> resident slots
        std::vector<int64_t> distinct;
        {
            std::unordered_map<int64_t, int> seen;
            for (int i = 0; i < n; ++i)
                if (seen.emplace(seeds[i], 1).second) distinct.push_back(seeds[i]);
        }
        ensure_seeds(distinct, S.stream);
        // chunks and the union of their 3x3 biome neighbourhoods (per seed)
        std::unordered_map<uint64_t, int> bindex;
        bindex.reserve(static_cast<size_t>(n) * 4);
        int nbk = 0;
        for (int i = 0; i < n; ++i) {
            const int slot = resident.at(seeds[i]);
            S.h_ch[i] = ChunkRef{xz[2 * i], xz[2 * i + 1], slot, 0};
            for (int dz = -1; dz <= 1; ++dz)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int bx = xz[2 * i] + dx, bz = xz[2 * i + 1] + dz;
                    const uint64_t key = (static_cast<uint64_t>(slot) << 48) ^ (static_cast<uint64_t>(static_cast<uint32_t>(bx) & 0xFFFFFF) << 24) ^
                                         static_cast<uint64_t>(static_cast<uint32_t>(bz) & 0xFFFFFF);
                    auto it = bindex.find(key);
                    if (it == bindex.end()) {
                        it = bindex.emplace(key, nbk).first;
                        S.h_bch[nbk++] = ChunkRef{bx, bz, slot, 0};
                    }
                    S.h_nb[static_cast<size_t>(i) * 9 + (dz + 1) * 3 + dx + 1] = it->second;
                }
        }
        S.ensure_biome_capacity(nbk);
        hipStream_t st = S.stream;
        hip_check(hipEventRecord(S.ev[0], st), "event");
        hip_check(hipMemcpyAsync(S.d_ch, S.h_ch, sizeof(ChunkRef) * n, hipMemcpyHostToDevice, st), "upload chunks");
        hip_check(hipMemcpyAsync(S.d_bch, S.h_bch, sizeof(ChunkRef) * nbk, hipMemcpyHostToDevice, st), "upload biome chunks");
        hip_check(hipMemcpyAsync(S.d_nb, S.h_nb, sizeof(int) * n * 9, hipMemcpyHostToDevice, st), "upload neighbours");
        hip_check(hipEventRecord(S.ev[1], st), "event");
        const SeedDev* sd = d_seeds;
        // biomes
        timed(S, kKBiomeColumns, [&] { k_biome_columns<<<grid_of(static_cast<size_t>(nbk) * 16, W(kKBiomeColumns)), W(kKBiomeColumns), 0, st>>>(sd, S.d_bch, nbk, S.d_bcol); });
        timed(S, kKBiomeQuarts, [&] {
            k_biome_targets<<<grid_of(static_cast<size_t>(nbk) * kQPer, 256), 256, 0, st>>>(sd, S.d_bch, nbk, S.d_bcol, S.d_targets);
            if (d_treef) {
                if (vary_groups > 0) {
                    k_biome_groups<<<grid_of(static_cast<size_t>(nbk) * 16 * vary_groups, W(kKBiomeQuarts)), W(kKBiomeQuarts), 0, st>>>(
                        nbk, vary_groups, S.d_targets, S.d_lstar, S.d_mind, S.d_gres);
                    k_biome_quart_final<<<grid_of(static_cast<size_t>(nbk) * kQPer, 256), 256, 0, st>>>(nbk, vary_groups, S.d_targets, S.d_gres, S.d_lstar,
                                                                                                         S.d_mind);
                } else {
                    k_biome_search_cols<<<grid_of(static_cast<size_t>(nbk) * 16, W(kKBiomeQuarts)), W(kKBiomeQuarts), 0, st>>>(
                        nbk, S.d_targets, S.d_lstar, S.d_mind);
                }
                k_biome_search_wide<<<grid_of(static_cast<size_t>(nbk) * kQPer, 256), 256, 0, st>>>(nbk, S.d_targets, S.d_lstar, S.d_mind);
            } else {
                k_biome_search_scalar<<<grid_of(static_cast<size_t>(nbk) * 16 * biome_runs(run_len), W(kKBiomeQuarts)), W(kKBiomeQuarts), 0, st>>>(
                    nbk, S.d_targets, S.d_lstar, S.d_mind, run_len);
            }
        });
        timed(S, kKBiomeTies, [&] { k_biome_ties<<<grid_of(static_cast<size_t>(nbk) * kQPer, W(kKBiomeTies)), W(kKBiomeTies), 0, st>>>(nbk, S.d_targets, S.d_lstar, S.d_mind, S.d_tie); });
        timed(S, kKBiomeChain, [&] { k_biome_chain<<<grid_of(static_cast<size_t>(nbk) * 64, W(kKBiomeChain)), W(kKBiomeChain), 0, st>>>(nbk, S.d_targets, S.d_lstar, S.d_mind, S.d_tie, S.d_bb, d_stats); });
        k_biome_masks<<<grid_of(static_cast<size_t>(nbk) * 16, 256), 256, 0, st>>>(nbk, S.d_bb, S.d_bmask);
        quarts += static_cast<uint64_t>(nbk) * kQPer;
        hip_check(hipGetLastError(), "biome kernels");
        hip_check(hipEventRecord(S.ev[2], st), "event");
        // density corners
        const int ck = n * kCols;
        const KeyCodec KC = codec_for(S.h_ch, n, 0, 4, 0, 4);
        const KeyCodec KG = codec_for(S.h_ch, n, -16, 10, -8, 10);
        k_col_keys<<<grid_of(ck, 256), 256, 0, st>>>(KC, S.d_ch, n, S.d_ckeys);
        timed(S, kKDedupeCols, [&] { S.dedupe->run(S.d_ckeys, ck, KC.bits, S.d_ucols, S.d_colmap, S.d_counts, st); });
        timed(S, kKColumns, [&] { k_columns<<<grid_of(ck, W(kKColumns)), W(kKColumns), 0, st>>>(KC, sd, ck, S.d_ucols, S.d_counts, S.d_colv); });
        timed(S, kKCorners, [&] { k_corners<<<grid_of(ck, kCornerCols), 256, 0, st>>>(KC, sd, ck, S.d_ucols, S.d_counts, S.d_colv, S.d_corners); });
        hip_check(hipGetLastError(), "noise kernels");
        hip_check(hipEventRecord(S.ev[3], st), "event");
        // surface-level grid and aquifer cells
        const int gk = n * kGridKeys;
        k_grid_keys<<<grid_of(gk, 256), 256, 0, st>>>(KG, S.d_ch, n, S.d_gkeys);
        timed(S, kKDedupeGrid, [&] { S.dedupe->run(S.d_gkeys, gk, KG.bits, S.d_ug, S.d_gmap, S.d_counts + 1, st); });
        timed(S, kKFts, [&] { k_fts<<<grid_of(gk, W(kKFts)), W(kKFts), 0, st>>>(KG, sd, gk, S.d_ug, S.d_counts + 1, S.d_gv); });
        timed(S, kKGridScatter, [&] { k_grid_scatter<<<grid_of(gk, 256), 256, 0, st>>>(n, S.d_gmap, S.d_gv, S.d_grid, S.d_prelim); });
        k_aq_skip<<<grid_of(n, 256), 256, 0, st>>>(n, S.d_grid, S.d_skip);
        timed(S, kKAqCells, [&] { k_aq_cells<<<grid_of(static_cast<size_t>(n) * kAqCells, W(kKAqCells)), W(kKAqCells), 0, st>>>(sd, S.d_ch, n, S.d_grid, S.d_cloc, S.d_cfl); });
        hip_check(hipGetLastError(), "aquifer kernels");
        hip_check(hipEventRecord(S.ev[4], st), "event");
        // fill
        uint8_t* blocks = to_device ? terrain : S.d_blocks;
        FillArgs F{sd, S.d_ch, S.d_corners, S.d_colmap, S.d_cloc, S.d_cfl, S.d_skip, blocks, nullptr, nullptr};
        FillArgs FH = F;
        FH.h0 = S.d_h0;
        FH.clsw = S.d_clsw;
        timed(S, kKHeights, [&] { k_heights_init<<<grid_of(static_cast<size_t>(n) * 256, W(kKHeights)), W(kKHeights), 0, st>>>(n, S.d_h0, S.d_counts + 2); });
        timed(S, kKFill, [&] { k_fill<<<grid_of(static_cast<size_t>(n) * 256 * kCellsY, W(kKFill)), W(kKFill), fill_lds_bytes(W(kKFill)), st>>>(FH, n); });
        hip_check(hipGetLastError(), "fill kernel");
        hip_check(hipEventRecord(S.ev[5], st), "event");
        if (aquifer) k_copy16<<<grid_of(static_cast<size_t>(n) * kPer / 16, 256), 256, 0, st>>>(static_cast<size_t>(n) * kPer / 16,
                                                                                               reinterpret_cast<const uint4*>(blocks),
                                                                                               reinterpret_cast<uint4*>(S.d_aq));
        hip_check(hipEventRecord(S.ev[6], st), "event");
        // surface
        SurfArgs A{sd, nullptr, S.d_ch, blocks, S.d_nb, S.d_bb, S.d_corners, S.d_colmap, S.d_prelim, S.d_colrec, S.d_ore, S.d_bmask, S.d_clsw};
        if (kOreRows > 0)
            timed(S, kKOre, [&] { k_ore<<<grid_of(static_cast<size_t>(n) * 256, W(kKOre)), W(kKOre), 0, st>>>(A, n); });
        timed(S, kKSurfaceCols, [&] { k_surface_cols<<<grid_of(static_cast<size_t>(n) * 256, W(kKSurfaceCols)), W(kKSurfaceCols), 0, st>>>(sd, S.d_ch, n, S.d_colrec, S.d_nb,
                                                                                                                                       S.d_bmask); });
        timed(S, kKSurface1, [&] { k_surface1<<<grid_of(static_cast<size_t>(n) * 256, 64), 64, 0, st>>>(A, n, S.d_h0, S.d_h1, S.d_h2); });
        timed(S, kKSurface2, [&] { k_surface2<<<grid_of(static_cast<size_t>(n) * 256, 64), 64, 0, st>>>(A, n, S.d_h0, S.d_h1, S.d_h2, S.d_hf); });
        k_surface_check<<<n, 256, 0, st>>>(n, S.d_h2, S.d_hf, S.d_flags, S.d_counts + 2, d_stats);
        timed(S, kKSurfaceExact, [&] { k_surface_exact<<<64, 256, 0, st>>>(F, A, S.d_flags, S.d_counts + 2); });
        uint8_t* bout = to_device ? biomes : S.d_bout;
        if (bout) k_gather_biomes<<<grid_of(static_cast<size_t>(n) * kQPer, 256), 256, 0, st>>>(n, S.d_nb, S.d_bb, bout);
        hip_check(hipGetLastError(), "surface kernels");
        hip_check(hipEventRecord(S.ev[7], st), "event");
        if (!to_device) {
            hip_check(hipMemcpyAsync(terrain, S.d_blocks, static_cast<size_t>(n) * kPer, hipMemcpyDeviceToHost, st), "download terrain");
            if (biomes) hip_check(hipMemcpyAsync(biomes, S.d_bout, static_cast<size_t>(n) * kQPer, hipMemcpyDeviceToHost, st), "download biomes");
            if (aquifer) hip_check(hipMemcpyAsync(aquifer, S.d_aq, static_cast<size_t>(n) * kPer, hipMemcpyDeviceToHost, st), "download aquifer");
        }
        hip_check(hipEventRecord(S.ev[8], st), "event");
        S.pending = true;
    }

    // Biome-only queries, in groups of at most kQueryGroup blocks and seed_cap distinct seeds.
    static constexpr int kQueryGroup = 1 << 20;
    int32_t *d_qxyz = nullptr, *d_qslot = nullptr, *d_qbad = nullptr;
    uint8_t* d_qout = nullptr;
    void biome_query(const int64_t* seeds, const int32_t* xyz, int count, uint8_t* out) {
        if (!d_treef) throw Error(MCT_ECAPACITY, "biome queries need the float climate tree (a bound of this model exceeds 2^23)");
        collect_all(nullptr);
        Slot& S = *slots[0];
        if (!d_qxyz) {
            hip_check(hipMalloc(&d_qxyz, sizeof(int32_t) * 3 * kQueryGroup), "malloc");
            hip_check(hipMalloc(&d_qslot, sizeof(int32_t) * kQueryGroup), "malloc");
            hip_check(hipMalloc(&d_qout, kQueryGroup), "malloc");
            hip_check(hipMalloc(&d_qbad, sizeof(int)), "malloc");
        }
        // Per group: local seed ids through a small open-addressing table, queries ordered by
        // seed (counting sort) so that a wave's lanes share a seed's tables, answers scattered
        // back to input order.
        const int cap = seed_cap;
        int tsize = 16;
        while (tsize < 4 * cap) tsize <<= 1;
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
                const int64_t sd = seeds[off + i];
                uint64_t h = static_cast<uint64_t>(sd) * 0x9E3779B97F4A7C15ull;
                size_t at = static_cast<size_t>(h >> 40) & static_cast<size_t>(tsize - 1);
                while (tval[at] >= 0 && tkey[at] != sd) at = (at + 1) & static_cast<size_t>(tsize - 1);
                if (tval[at] < 0) {
                    if (static_cast<int>(distinct.size()) == cap) { n = i; break; }
                    tkey[at] = sd;
                    tval[at] = static_cast<int32_t>(distinct.size());
                    distinct.push_back(sd);
                }
                lid[i] = tval[at];
            }
            ensure_seeds(distinct, S.stream);
            const int nd = static_cast<int>(distinct.size());
            cnt.assign(static_cast<size_t>(nd) + 1, 0);
            for (int i = 0; i < n; ++i) ++cnt[lid[i] + 1];
            for (int d = 0; d < nd; ++d) cnt[d + 1] += cnt[d];
            order.resize(static_cast<size_t>(n));
            slot.resize(static_cast<size_t>(n));
            sxyz.resize(static_cast<size_t>(n) * 3);
            for (int i = 0; i < n; ++i) order[cnt[lid[i]]++] = i;
            std::vector<int32_t> dslot(static_cast<size_t>(nd));
            for (int d = 0; d < nd; ++d) dslot[d] = resident.at(distinct[d]);
            for (int k = 0; k < n; ++k) {
                const size_t i = static_cast<size_t>(off) + static_cast<size_t>(order[k]);
                slot[k] = dslot[lid[order[k]]];
                sxyz[3 * static_cast<size_t>(k)] = xyz[3 * i];
                sxyz[3 * static_cast<size_t>(k) + 1] = xyz[3 * i + 1];
                sxyz[3 * static_cast<size_t>(k) + 2] = xyz[3 * i + 2];
            }
            hip_check(hipMemcpyAsync(d_qxyz, sxyz.data(), sizeof(int32_t) * 3 * n, hipMemcpyHostToDevice, S.stream), "upload");
            hip_check(hipMemcpyAsync(d_qslot, slot.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, S.stream), "upload");
            hip_check(hipMemsetAsync(d_qbad, 0, sizeof(int), S.stream), "memset");
            k_biome_query<<<grid_of(n, 256), 256, 0, S.stream>>>(d_seeds, d_qslot, d_qxyz, n, d_qout, d_qbad);
            hip_check(hipGetLastError(), "query kernel");
            int bad = 0;
            sout.resize(static_cast<size_t>(n));
            hip_check(hipMemcpyAsync(sout.data(), d_qout, n, hipMemcpyDeviceToHost, S.stream), "download");
            hip_check(hipMemcpyAsync(&bad, d_qbad, sizeof(int), hipMemcpyDeviceToHost, S.stream), "download");
            hip_check(hipStreamSynchronize(S.stream), "sync");
            if (bad) throw Error(MCT_ECAPACITY, std::to_string(bad) + " queries have climate targets beyond 2^23 (not served by the float tree)");
            for (int k = 0; k < n; ++k) out[off + order[k]] = sout[k];
            off += n;
        }
    }

    // Kernel groups whose workgroup size is a launch parameter, and the sizes tried.
    static bool tunable(int k) {
        switch (k) {
            case kKBiomeColumns: case kKBiomeQuarts: case kKBiomeTies: case kKBiomeChain: case kKColumns: case kKFts: case kKAqCells:
            case kKFill: case kKHeights: case kKSurfaceCols: case kKOre: return true;
            default: return false;
        }
    }

    // Times each launch shape on a sample (the first max_batch chunks of it, outputs in HBM):
    // per tunable kernel group, the workgroup sizes 64 / 128 / 256, and for the R-tree search the
    // quarts per thread; keeps the fastest of each (coordinate descent, one knob at a time, the
    // group's own HIP-event time, best of two timed passes after a warm pass).
    void autotune(const int64_t* seeds, const int32_t* xz, int count, mct_tuning* out) {
        count = std::min(count, max_batch);
        collect_all(nullptr);
        static_assert(kKCount <= 24, "mct_tuning holds 24 kernel groups");
        uint8_t* d_t = nullptr;
        hip_check(hipMalloc(&d_t, static_cast<size_t>(count) * kPer), "malloc");
        const bool was = profile;
        profile = true;
        double best_ms[kKCount] = {};
        auto group_ms = [&](int k) {
            double best = 1e300;
            generate(seeds, xz, count, d_t, nullptr, nullptr, nullptr, true);
            for (int rep = 0; rep < 2; ++rep) {
                for (int g = 0; g < kKCount; ++g) kernel_ms[g] = 0.0, kernel_launches[g] = 0;
                generate(seeds, xz, count, d_t, nullptr, nullptr, nullptr, true);
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
        } catch (...) {
            profile = was;
            (void)hipFree(d_t);
            throw;
        }
        profile = was;
        for (int g = 0; g < kKCount; ++g) kernel_ms[g] = 0.0, kernel_launches[g] = 0;
        hip_check(hipFree(d_t), "free");
        if (out) {
            get_tuning(out);
            for (int g = 0; g < kKCount; ++g) out->ms[g] = best_ms[g];
        }
    }

    void get_tuning(mct_tuning* out) const {
        *out = mct_tuning{};
        for (int k = 0; k < kKCount; ++k) out->workgroup[k] = tunable(k) ? W(k) : 0;
        out->run_len = run_len;
    }

    void set_tuning(const mct_tuning& t) {
        for (int k = 0; k < kKCount; ++k) {
            const int v = t.workgroup[k];
            if (tunable(k) ? !(v == 0 || v == 64 || v == 128 || v == 256) : v != 0)
                throw Error(MCT_EINVAL, std::string("workgroup size of ") + kKernelNames[k] + " must be " + (tunable(k) ? "0, 64, 128 or 256" : "0"));
        }
        if (t.run_len < 1 || t.run_len > kQH) throw Error(MCT_EINVAL, "run_len must be 1.." + std::to_string(kQH));
        collect_all(nullptr);
        for (int k = 0; k < kKCount; ++k) wg[k] = t.workgroup[k];
        run_len = t.run_len;
    }

    void generate(const int64_t* seeds, const int32_t* xz, int count, uint8_t* terrain, uint8_t* biomes, uint8_t* aquifer, mct_timing* tm,
                  bool to_device, bool core = false, float* density = nullptr) {
        const auto w0 = std::chrono::steady_clock::now();
        collect_all(nullptr);
        done_ms.assign(static_cast<size_t>((count + max_batch - 1) / max_batch), 0.0);
        hip_check(hipEventRecord(t0, slots[0]->stream), "event");
        int batch = 0;
        for (int off = 0; off < count; ++batch) {
            int n = std::min(max_batch, count - off);
            {   // at most seed_cap distinct seeds per batch
                std::unordered_map<int64_t, int> seen;
                for (int i = 0; i < n; ++i) {
                    if (seen.emplace(seeds[off + i], 1).second && static_cast<int>(seen.size()) > seed_cap) { n = i; break; }
                }
            }
            if (batch >= static_cast<int>(done_ms.size())) done_ms.push_back(0.0);
            Slot& S = *slots[next_slot];
            next_slot = (next_slot + 1) % slots.size();
            collect(S, tm);
            S.batch = batch;
            if (core)
                run_batch_core(S, seeds + off, xz + 2 * static_cast<size_t>(off), n, terrain ? terrain + static_cast<size_t>(off) * kPer : nullptr,
                               density ? density + static_cast<size_t>(off) * kPer : nullptr, to_device);
            else
                run_batch(S, seeds + off, xz + 2 * static_cast<size_t>(off), n, terrain + static_cast<size_t>(off) * kPer,
                          biomes ? biomes + static_cast<size_t>(off) * kQPer : nullptr, aquifer ? aquifer + static_cast<size_t>(off) * kPer : nullptr,
                          to_device);
            off += n;
        }
        collect_all(tm);
        if (tm) tm->wall_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    }
};

}  // namespace

#include "server.h"

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

}  // namespace

struct mct_generator {
    std::unique_ptr<Generator> g;
};

struct mct_server {
    std::unique_ptr<Server> s;
};

namespace {
void check_model_files(const char* ir_json, const char* terrain_json) {
    if (!ir_json || !terrain_json) throw Error(MCT_EINVAL, "null path");
    if (file_fnv1a(ir_json) != mctgen::kIrHash) throw Error(MCT_EMODEL, "the IR file differs from the one this library was generated from (rebuild)");
    if (file_fnv1a(terrain_json) != mctgen::kTerrainHash)
        throw Error(MCT_EMODEL, "the terrain model file differs from the one this library was generated from (rebuild)");
}
}  // namespace

extern "C" {

const char* mct_last_error(void) { return g_err.c_str(); }

int mct_create(const char* ir_json, const char* terrain_json, int max_batch_chunks, int pipeline_slots, int seed_capacity, mct_generator** out) {
    return guard([&] {
        if (!out) throw Error(MCT_EINVAL, "out is null");
        *out = nullptr;
        if (!ir_json || !terrain_json) throw Error(MCT_EINVAL, "null path");
        if (max_batch_chunks < 1 || max_batch_chunks > (1 << 20)) throw Error(MCT_EINVAL, "max_batch_chunks out of range");
        if (pipeline_slots < 1 || pipeline_slots > 4) throw Error(MCT_EINVAL, "pipeline_slots must be 1..4");
        if (seed_capacity == 0) seed_capacity = 64;
        if (seed_capacity < 1 || seed_capacity > 4096) throw Error(MCT_EINVAL, "seed_capacity must be 1..4096");
        check_model_files(ir_json, terrain_json);
        int dev = 0;
        hip_check(hipGetDevice(&dev), "get device");
        auto h = std::make_unique<mct_generator>();
        h->g = std::make_unique<Generator>(max_batch_chunks, pipeline_slots, seed_capacity);
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
        g->g->biome_query(seeds, xyz, count, biomes);
    });
}

int mct_zoom_selftest(int64_t world_seed, int64_t count, int64_t out[3]) {
    return guard([&] {
        if (!out || count < 0) throw Error(MCT_EINVAL, "bad arguments");
        const int64_t zoom_seed = static_cast<int64_t>(sha256_long_le(world_seed));
        unsigned long long* d = nullptr;
        hip_check(hipMalloc(&d, sizeof(unsigned long long) * 3), "malloc");
        hip_check(hipMemset(d, 0, sizeof(unsigned long long) * 3), "memset");
        const int64_t columns = (count + kHeight - 1) / kHeight;
        for (int64_t c0 = 0; c0 < columns; c0 += (1 << 24)) {
            const int nc = static_cast<int>(std::min<int64_t>(1 << 24, columns - c0));
            k_zoom_selftest<<<grid_of(nc, 256), 256>>>(zoom_seed, 0x5A4F4F4Dull + static_cast<uint64_t>(c0), nc, d);
            hip_check(hipGetLastError(), "zoom selftest kernel");
        }
        unsigned long long h[3];
        hip_check(hipMemcpy(h, d, sizeof h, hipMemcpyDeviceToHost), "download");
        (void)hipFree(d);
        out[0] = static_cast<int64_t>(h[0]);
        out[1] = static_cast<int64_t>(h[1]);
        out[2] = static_cast<int64_t>(h[2]);
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
        g->g->collect_all(nullptr);
        g->g->current_seed = seed;
    });
}

int mct_generate(mct_generator* g, const int32_t* chunk_xz, int count, uint8_t* terrain, uint8_t* biomes, uint8_t* aquifer, mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!chunk_xz || !terrain))) throw Error(MCT_EINVAL, "bad arguments");
        std::vector<int64_t> seeds(static_cast<size_t>(count), g->g->current_seed);
        g->g->generate(seeds.data(), chunk_xz, count, terrain, biomes, aquifer, timing, false);
    });
}

int mct_generate_multi(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, uint8_t* terrain, uint8_t* biomes, uint8_t* aquifer,
                       mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!seeds || !chunk_xz || !terrain))) throw Error(MCT_EINVAL, "bad arguments");
        g->g->generate(seeds, chunk_xz, count, terrain, biomes, aquifer, timing, false);
    });
}

int mct_generate_device(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, uint8_t* d_terrain, uint8_t* d_biomes,
                        mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!seeds || !chunk_xz || !d_terrain))) throw Error(MCT_EINVAL, "bad arguments");
        g->g->generate(seeds, chunk_xz, count, d_terrain, d_biomes, nullptr, timing, true);
    });
}

int mct_generate_core(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, float* density, uint8_t* fill, mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!seeds || !chunk_xz || (!density && !fill)))) throw Error(MCT_EINVAL, "bad arguments");
        g->g->generate(seeds, chunk_xz, count, fill, nullptr, nullptr, timing, false, true, density);
    });
}

int mct_generate_core_device(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, float* d_density, uint8_t* d_fill,
                             mct_timing* timing) {
    return guard([&] {
        if (!g || count < 0 || (count > 0 && (!seeds || !chunk_xz || (!d_density && !d_fill)))) throw Error(MCT_EINVAL, "bad arguments");
        g->g->generate(seeds, chunk_xz, count, d_fill, nullptr, nullptr, timing, true, true, d_density);
    });
}

int mct_autotune(mct_generator* g, const int64_t* seeds, const int32_t* chunk_xz, int count, mct_tuning* out) {
    return guard([&] {
        if (!g || count < 1 || !seeds || !chunk_xz) throw Error(MCT_EINVAL, "bad arguments");
        g->g->autotune(seeds, chunk_xz, count, out);
    });
}

int mct_get_tuning(mct_generator* g, mct_tuning* out) {
    return guard([&] {
        if (!g || !out) throw Error(MCT_EINVAL, "bad arguments");
        g->g->get_tuning(out);
    });
}

int mct_set_tuning(mct_generator* g, const mct_tuning* t) {
    return guard([&] {
        if (!g || !t) throw Error(MCT_EINVAL, "bad arguments");
        g->g->set_tuning(*t);
    });
}

int mct_biome_tie_stats(mct_generator* g, int64_t out[4]) {
    return guard([&] {
        if (!g || !out) throw Error(MCT_EINVAL, "bad arguments");
        g->g->collect_all(nullptr);
        unsigned long long s[4];
        hip_check(hipMemcpy(s, g->g->d_stats, sizeof s, hipMemcpyDeviceToHost), "download stats");
        out[0] = static_cast<int64_t>(g->g->quarts);
        out[1] = static_cast<int64_t>(s[1]);
        out[2] = static_cast<int64_t>(s[2]);
        out[3] = static_cast<int64_t>(s[3]);
    });
}

int mct_last_batch_done_ms(mct_generator* g, double* out, int cap, int* n) {
    return guard([&] {
        if (!g || !n) throw Error(MCT_EINVAL, "bad arguments");
        const auto& d = g->g->done_ms;
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
                k_topdown<<<grid_of(static_cast<size_t>(nc) * 256, 256), 256>>>(src, nc, d_out);
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

int mct_profile(mct_generator* g, int enable) {
    return guard([&] {
        if (!g) throw Error(MCT_EINVAL, "null generator");
        g->g->collect_all(nullptr);
        g->g->profile = enable != 0;
        for (int k = 0; k < kKCount; ++k) {
            g->g->kernel_ms[k] = 0;
            g->g->kernel_launches[k] = 0;
        }
    });
}

int mct_kernel_times(mct_generator* g, double* ms, int64_t* launches, int cap, int* n) {
    return guard([&] {
        if (!g || !n) throw Error(MCT_EINVAL, "bad arguments");
        g->g->collect_all(nullptr);
        *n = kKCount;
        for (int k = 0; k < cap && k < kKCount; ++k) {
            if (ms) ms[k] = g->g->kernel_ms[k];
            if (launches) launches[k] = g->g->kernel_launches[k];
        }
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
