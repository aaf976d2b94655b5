// Random sources of the TERRAIN step, host and device:
//   Xoroshiro128++ (Blackman & Vigna) as vanilla's XoroshiroRandomSource with its
//   positional factory, Lemire's bounded nextInt, RandomSupport's Stafford-13 mix and
//   128-bit seed upgrade, Mth.getSeed, BiomeManager's LCG fiddle, and SHA-256 (FIPS 180-4)
//   for the obfuscated zoom seed. MD5 (RFC 1321) of names is host-only: names are
//   seed-independent, so their digests are computed once.
// Reference semantics: cpu/src/rng.cpp:19-210, cpu/src/biome.cpp:14-70 and 241-289,
// cuda/include/mccuda_cuda/terrain_device.h:12-68.
#pragma once

#include "jmath.h"

namespace mct {

struct Xoro {
    uint64_t lo, hi;
};

MCT_HD MCT_FI uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// Xoroshiro128PlusPlus(lo, hi): an all-zero state is replaced by the golden/silver ratios.
MCT_HD MCT_FI Xoro xoro(uint64_t lo, uint64_t hi) {
    if ((lo | hi) == 0) return {0x9E3779B97F4A7C15ull, 0x6A09E667F3BCC909ull};
    return {lo, hi};
}

MCT_HD MCT_FI uint64_t next64(Xoro& s) {
    const uint64_t s0 = s.lo;
    uint64_t s1 = s.hi;
    const uint64_t r = rotl(s0 + s1, 17) + s0;
    s1 ^= s0;
    s.lo = rotl(s0, 49) ^ s1 ^ (s1 << 21);
    s.hi = rotl(s1, 28);
    return r;
}

// nextInt(bound): Lemire multiply-shift with rejection below (-bound) mod bound.
MCT_HD MCT_FI int32_t next_int(Xoro& s, int32_t bound) {
    uint64_t m = static_cast<uint64_t>(static_cast<uint32_t>(next64(s))) * static_cast<uint64_t>(static_cast<uint32_t>(bound));
    if ((m & 0xFFFFFFFFull) < static_cast<uint64_t>(static_cast<uint32_t>(bound))) {
        const uint32_t ub = static_cast<uint32_t>(bound);
        const uint32_t threshold = (0u - ub) % ub;
        while ((m & 0xFFFFFFFFull) < threshold) m = static_cast<uint64_t>(static_cast<uint32_t>(next64(s))) * static_cast<uint64_t>(ub);
    }
    return static_cast<int32_t>(m >> 32);
}

// nextFloat: top 24 bits times 2^-24 (float); nextDouble: top 53 bits times 2^-53 (double).
MCT_HD MCT_FI float next_float(Xoro& s) { return static_cast<float>(static_cast<int64_t>(next64(s) >> 40)) * 0x1p-24f; }
MCT_HD MCT_FI double next_double(Xoro& s) { return static_cast<double>(static_cast<int64_t>(next64(s) >> 11)) * 0x1p-53; }
MCT_HD MCT_FI bool next_bool(Xoro& s) { return (next64(s) & 1ull) != 0; }

MCT_HD MCT_FI uint64_t stafford13(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// XoroshiroRandomSource(long seed): RandomSupport.upgradeSeedTo128bit.
MCT_HD MCT_FI Xoro xoro_from_seed(int64_t seed) {
    const uint64_t lo = static_cast<uint64_t>(seed) ^ 0x6A09E667F3BCC909ull;
    const uint64_t hi = lo + 0x9E3779B97F4A7C15ull;
    return xoro(stafford13(lo), stafford13(hi));
}

// forkPositional(): the next two longs are the factory's seed.
struct Positional {
    uint64_t lo, hi;
};
MCT_HD MCT_FI Positional fork_positional(Xoro& s) {
    const uint64_t lo = next64(s);
    const uint64_t hi = next64(s);
    return {lo, hi};
}

// fromHashOf(name): the name's MD5 (big-endian longs) xor the factory seed.
struct Digest {
    uint64_t lo, hi;
};
MCT_HD MCT_FI Xoro from_hash(const Positional& p, const Digest& d) { return xoro(d.lo ^ p.lo, d.hi ^ p.hi); }

// Mth.getSeed(x, y, z)
MCT_HD MCT_FI int64_t pos_seed(int32_t x, int32_t y, int32_t z) {
    const int64_t xm = static_cast<int64_t>(imul(x, 3129871));
    uint64_t s = static_cast<uint64_t>(xm) ^ (static_cast<uint64_t>(static_cast<int64_t>(z)) * 116129781ull) ^ static_cast<uint64_t>(static_cast<int64_t>(y));
    s = s * s * 42317861ull + s * 11ull;
    return static_cast<int64_t>(s) >> 16;
}

// positional.at(x, y, z)
MCT_HD MCT_FI Xoro at(const Positional& p, int32_t x, int32_t y, int32_t z) {
    return xoro(static_cast<uint64_t>(pos_seed(x, y, z)) ^ p.lo, p.hi);
}

// ---- BiomeManager zoom -------------------------------------------------------------------
// LinearCongruentialGenerator.next(rval, c) = rval * (rval * 6364136223846793005 + 1442695040888963407) + c
MCT_HD MCT_FI int64_t lcg(int64_t r, int64_t c) {
    const uint64_t u = static_cast<uint64_t>(r);
    return static_cast<int64_t>(u * (u * 6364136223846793005ull + 1442695040888963407ull) + static_cast<uint64_t>(c));
}

// getFiddle index: floorMod(rval >> 24, 1024)
MCT_HD MCT_FI int fiddle_index(int64_t r) {
    int64_t m = (r >> 24) % 1024;
    return static_cast<int>(m < 0 ? m + 1024 : m);
}

MCT_HD MCT_FI double fiddle_value(int m) { return (static_cast<double>(m) / 1024.0 - 0.5) * 0.9; }

// The three fiddle indices of a zoom corner (its LCG chain depends on the quart only).
MCT_HD MCT_FI void corner_fiddles(int64_t seed, int32_t qx, int32_t qy, int32_t qz, int& mx, int& my, int& mz) {
    int64_t r = seed;
    r = lcg(r, qx);
    r = lcg(r, qy);
    r = lcg(r, qz);
    r = lcg(r, qx);
    r = lcg(r, qy);
    r = lcg(r, qz);
    mx = fiddle_index(r);
    r = lcg(r, seed);
    my = fiddle_index(r);
    r = lcg(r, seed);
    mz = fiddle_index(r);
}

// (d + fiddle)^2 for one axis: d = (a & 3) / 4 on the near corner, minus 1 on the far one.
MCT_HD MCT_FI double zoom_sq(int frac4, bool far, int m) {
    double d = static_cast<double>(frac4) / 4.0;
    if (far) d = d - 1.0;
    const double v = fiddle_value(m);
    return (d + v) * (d + v);
}

// BiomeManager.getBiome's quart choice for block (x, y, z): the corner of the 8 around
// (x-2, y-2, z-2) >> 2 with the smallest fiddled distance (sz + sy) + sx, first wins ties.
MCT_HD MCT_FI void zoom_pick(int64_t seed, int32_t x, int32_t y, int32_t z, int32_t& qx, int32_t& qy, int32_t& qz) {
    const int32_t ax = x - 2, ay = y - 2, az = z - 2;
    const int32_t px = ax >> 2, py = ay >> 2, pz = az >> 2;
    const int fx = ax & 3, fy = ay & 3, fz = az & 3;
    int best = 0;
    double best_d = u2d(0x7FF0000000000000ull);
    for (int i = 0; i < 8; ++i) {
        const bool xf = (i & 4) != 0, yf = (i & 2) != 0, zf = (i & 1) != 0;
        int mx, my, mz;
        corner_fiddles(seed, xf ? px + 1 : px, yf ? py + 1 : py, zf ? pz + 1 : pz, mx, my, mz);
        const double d = (zoom_sq(fz, zf, mz) + zoom_sq(fy, yf, my)) + zoom_sq(fx, xf, mx);
        if (best_d > d) {
            best = i;
            best_d = d;
        }
    }
    qx = (best & 4) ? px + 1 : px;
    qy = (best & 2) ? py + 1 : py;
    qz = (best & 1) ? pz + 1 : pz;
}

// ---- SHA-256 of a long (Guava hashLong: little-endian bytes), first 8 digest bytes LE -------
MCT_HD inline uint64_t sha256_long_le(int64_t value) {
    const uint32_t K[64] = {
        0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u, 0xd807aa98u, 0x12835b01u,
        0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu,
        0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u,
        0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
        0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u, 0x1e376c08u,
        0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u, 0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u,
        0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};
    uint32_t w[64];
    // one block: 8 message bytes (little-endian long), 0x80, zeros, bit length 64 (big-endian)
    const uint64_t v = static_cast<uint64_t>(value);
    uint8_t b[8];
    for (int i = 0; i < 8; ++i) b[i] = static_cast<uint8_t>(v >> (8 * i));
    w[0] = (static_cast<uint32_t>(b[0]) << 24) | (static_cast<uint32_t>(b[1]) << 16) | (static_cast<uint32_t>(b[2]) << 8) | b[3];
    w[1] = (static_cast<uint32_t>(b[4]) << 24) | (static_cast<uint32_t>(b[5]) << 16) | (static_cast<uint32_t>(b[6]) << 8) | b[7];
    w[2] = 0x80000000u;
    for (int i = 3; i < 15; ++i) w[i] = 0;
    w[15] = 64;
    auto rr = [](uint32_t x, int n) { return (x >> n) | (x << (32 - n)); };
    for (int i = 16; i < 64; ++i) {
        const uint32_t s0 = rr(w[i - 15], 7) ^ rr(w[i - 15], 18) ^ (w[i - 15] >> 3);
        const uint32_t s1 = rr(w[i - 2], 17) ^ rr(w[i - 2], 19) ^ (w[i - 2] >> 10);
        w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t h[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au, 0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
    uint32_t a = h[0], bb = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; ++i) {
        const uint32_t S1 = rr(e, 6) ^ rr(e, 11) ^ rr(e, 25);
        const uint32_t ch = (e & f) ^ (~e & g);
        const uint32_t t1 = hh + S1 + ch + K[i] + w[i];
        const uint32_t S0 = rr(a, 2) ^ rr(a, 13) ^ rr(a, 22);
        const uint32_t mj = (a & bb) ^ (a & c) ^ (bb & c);
        const uint32_t t2 = S0 + mj;
        hh = g; g = f; f = e; e = d + t1; d = c; c = bb; bb = a; a = t1 + t2;
    }
    h[0] += a; h[1] += bb;
    // digest bytes 0..7 are h[0], h[1] big-endian; read them as a little-endian long
    uint64_t r = 0;
    for (int i = 7; i >= 0; --i) {
        const uint32_t word = h[i >> 2];
        const uint8_t byte = static_cast<uint8_t>(word >> (24 - 8 * (i & 3)));
        r = (r << 8) | byte;
    }
    return r;
}

}  // namespace mct
