// Device dedupe of 64-bit position keys: an LSD radix sort of (key, index) pairs, run heads,
// a scan of the heads and a scatter give the distinct keys in increasing order, the map from
// each input position to its key's rank, and the distinct count, all left on the device.
//
// The element count is read from device memory and every launch is sized for the capacity
// (tiles past the count exit), so the same launches serve any batch: the dedupe is part of the
// batch's captured graph. The number of 8-bit passes is the only host-side shape (the key
// width of the batch's codec).
#pragma once

#include "kernels.h"

namespace mct {

constexpr int kRsThreads = 256, kRsItems = 16, kRsTile = kRsThreads * kRsItems;
constexpr int kRsWaves = kRsThreads / 64, kRsRounds = kRsItems;   // per wave: kRsRounds rounds of 64 keys
constexpr int kScanTile = 2048;                                    // 256 threads x 8

__device__ MCT_FI int rs_tiles(int n) { return (n + kRsTile - 1) / kRsTile; }
__device__ MCT_FI int scan_tiles(int total) { return (total + kScanTile - 1) / kScanTile; }

// Per-tile digit histograms, digit-major: hist[d * tiles + tile].
__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint64_t* keys, const int* np, int shift, int* hist) {
    const int n = *np, nb = rs_tiles(n);
    if (static_cast<int>(blockIdx.x) >= nb) return;
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRsTile;
    for (int k = 0; k < kRsItems; ++k) {
        const int i = base + k * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of a long int array in three launches: per-tile sums, a scan of the tile
// sums, then each tile scanned in LDS from its offset (integer sums: any order is exact).
// total = 256 x (radix tiles of *np): the histogram of a pass.
__global__ __launch_bounds__(256) void k_scan_tiles(const int* v, const int* np, int* sums) {
    const int total = 256 * rs_tiles(*np);
    if (static_cast<int>(blockIdx.x) >= scan_tiles(total)) return;
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

// Exclusive scan of `total` ints in one workgroup (each thread a contiguous run). What is
// scanned: kind 0 the scan-tile sums of a pass's histogram, kind 1 the radix-tile sums of the
// head flags (then *total_out = their sum, the distinct count).
__global__ __launch_bounds__(1024) void k_scan_small(int* v, const int* np, int kind, int* total_out) {
    const int nb = rs_tiles(*np);
    const int total = kind == 0 ? scan_tiles(256 * nb) : nb;
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

__global__ __launch_bounds__(256) void k_scan_apply(int* v, const int* np, const int* offsets) {
    const int total = 256 * rs_tiles(*np);
    if (static_cast<int>(blockIdx.x) >= scan_tiles(total)) return;
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

__device__ MCT_FI unsigned long long match_digit(int d, bool valid) {
    unsigned long long m = __ballot(valid);
    for (int b = 0; b < 8; ++b) {
        const unsigned long long bb = __ballot(valid && ((d >> b) & 1));
        m &= ((d >> b) & 1) ? bb : ~bb;
    }
    return m;
}

// Stable scatter: wave w owns keys [base + w * 512, +512) in rounds of 64; ranks come from
// 64-lane ballots, per-wave digit counts from a first pass.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint64_t* kin, const int* vin, const int* np, int shift, const int* hist, uint64_t* kout,
                                                          int* vout) {
    const int n = *np, nb = rs_tiles(n);
    if (static_cast<int>(blockIdx.x) >= nb) return;
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
        int run = hist[d * nb + blockIdx.x];
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

__global__ __launch_bounds__(256) void k_heads(const uint64_t* sorted, const int* np, int* flag) {
    const int n = *np;
    for (int i = global_id(); i < n; i += global_threads()) flag[i] = (i == 0 || sorted[i] != sorted[i - 1]) ? 1 : 0;
}

// Per radix-tile sums of the head flags.
__global__ __launch_bounds__(256) void k_tile_sums(const int* flag, const int* np, int* sums) {
    const int n = *np;
    if (static_cast<int>(blockIdx.x) >= rs_tiles(n)) return;
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
__global__ __launch_bounds__(256) void k_unique_scatter(const uint64_t* sorted, const int* idx, const int* flag, const int* np, const int* tile_off,
                                                        uint64_t* unique, int* map, int* count, const int* total) {
    const int n = *np;
    if (static_cast<int>(blockIdx.x) >= rs_tiles(n)) return;
    __shared__ int s[kRsTile];
    const int base = blockIdx.x * kRsTile;
    for (int k = 0; k < kRsItems; ++k) {
        const int i = base + k * 256 + threadIdx.x;
        s[k * 256 + threadIdx.x] = i < n ? flag[i] : 0;
    }
    __syncthreads();
    // inclusive scan of the tile's flags: each thread scans kRsItems consecutive, then a
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

// ---- dense form -----------------------------------------------------------------------------
// For a batch whose key space [0, dense_size(K)) is small (contiguous regions: the codec
// compresses each axis to the batch's ranges), the same outputs without a sort: a flag per
// possible key, the exclusive scan of the flags is each present key's rank (the keys' order is
// the dense order), unique[rank] = the key, map[i] = the rank of key i. Six launches whatever the
// key width; the per-key work is two scattered int accesses into an array of dense_size ints.
// f holds at least dense_size(K) ints; tsum a tile sum per kDenseTile flags, plus one.
constexpr int kDenseTile = 2048;   // flags per scan tile: 256 threads x 8

__device__ MCT_FI int dense_n(const KeyCodec* K) { return static_cast<int>(dense_size(*K)); }

__global__ __launch_bounds__(256) void k_dense_clear(const KeyCodec* K, int* f) {
    const int n = dense_n(K);
    for (int i = global_id(); i < n; i += global_threads()) f[i] = 0;
}

__global__ __launch_bounds__(256) void k_dense_mark(const uint64_t* keys, const int* np, const KeyCodec* K, int* f) {
    const int n = *np;
    for (int i = global_id(); i < n; i += global_threads()) f[dense_index(*K, keys[i])] = 1;
}

// Flags set per tile of kDenseTile (workgroups stride over the tiles: any grid).
__global__ __launch_bounds__(256) void k_dense_tiles(const KeyCodec* K, const int* f, int* tsum) {
    const int n = dense_n(K), tiles = (n + kDenseTile - 1) / kDenseTile;
    __shared__ int part[256];
    for (int tile = static_cast<int>(blockIdx.x); tile < tiles; tile += static_cast<int>(gridDim.x)) {
        const int base = tile * kDenseTile;
        int a = 0;
        for (int j = 0; j < kDenseTile / 256; ++j) {
            const int i = base + j * 256 + static_cast<int>(threadIdx.x);
            if (i < n) a += f[i];
        }
        part[threadIdx.x] = a;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (static_cast<int>(threadIdx.x) < w) part[threadIdx.x] += part[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) tsum[tile] = part[0];
        __syncthreads();
    }
}

// Exclusive scan of the tile sums in one workgroup; *count = the distinct keys.
__global__ __launch_bounds__(1024) void k_dense_offsets(const KeyCodec* K, int* tsum, int* count) {
    const int tiles = (dense_n(K) + kDenseTile - 1) / kDenseTile;
    __shared__ int part[1024];
    const int t = static_cast<int>(threadIdx.x);
    const int per = (tiles + 1023) / 1024;
    const int b = t * per, e = b + per < tiles ? b + per : tiles;
    int sum = 0;
    for (int i = b; i < e; ++i) sum += tsum[i];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int a = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += a;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int i = b; i < e; ++i) {
        const int x = tsum[i];
        tsum[i] = run;
        run += x;
    }
    if (t == 1023) *count = part[1023];
}

// Per tile: each set flag becomes its key's rank (the tile's offset plus the set flags before
// it in the tile) and unique[rank] the key (workgroups stride over the tiles: any grid).
__global__ __launch_bounds__(256) void k_dense_ranks(const KeyCodec* K, int* f, const int* tsum, uint64_t* unique) {
    const int n = dense_n(K), tiles = (n + kDenseTile - 1) / kDenseTile;
    constexpr int kPer = kDenseTile / 256;
    __shared__ int ts[256];
    const int t = static_cast<int>(threadIdx.x);
    for (int tile = static_cast<int>(blockIdx.x); tile < tiles; tile += static_cast<int>(gridDim.x)) {
        const int base = tile * kDenseTile + t * kPer;   // a thread's kPer consecutive flags
        int v[kPer];
        int a = 0;
        for (int k = 0; k < kPer; ++k) {
            v[k] = base + k < n ? f[base + k] : 0;
            a += v[k];
        }
        ts[t] = a;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int x = t >= off ? ts[t - off] : 0;
            __syncthreads();
            ts[t] += x;
            __syncthreads();
        }
        int run = tsum[tile] + ts[t] - a;
        for (int k = 0; k < kPer; ++k) {
            if (v[k]) {
                f[base + k] = run;
                unique[run] = dense_key(*K, static_cast<uint32_t>(base + k));
                ++run;
            }
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_dense_map(const uint64_t* keys, const int* np, const KeyCodec* K, const int* f, int* map) {
    const int n = *np;
    for (int i = global_id(); i < n; i += global_threads()) map[i] = f[dense_index(*K, keys[i])];
}

}  // namespace mct
