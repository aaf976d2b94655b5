// Oracle: Java arithmetic, random sources, hashes, gradient noise and the JSON reader.
// Restates cpu/include/mccuda/javamath.h, cpu/src/rng.cpp, cpu/include/mccuda/noise_eval.h
// and the vanilla classes they mirror (Mth, XoroshiroRandomSource, LegacyRandomSource,
// RandomSupport, PerlinNoise, SmearedPerlinNoise, GradientNoise).
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "oracle.h"

namespace orc {

// ---- Java arithmetic ---------------------------------------------------------------------------
float java_min(float a, float b) {
    if (std::isnan(a)) return a;
    if (a == 0.0f && b == 0.0f && std::signbit(b)) return b;
    return a <= b ? a : b;
}
float java_max(float a, float b) {
    if (std::isnan(a)) return a;
    if (a == 0.0f && b == 0.0f && std::signbit(a)) return b;
    return a >= b ? a : b;
}
double java_min(double a, double b) {
    if (std::isnan(a)) return a;
    if (a == 0.0 && b == 0.0 && std::signbit(b)) return b;
    return a <= b ? a : b;
}
double java_max(double a, double b) {
    if (std::isnan(a)) return a;
    if (a == 0.0 && b == 0.0 && std::signbit(a)) return b;
    return a >= b ? a : b;
}
float lerp3(float a1, float a2, float a3, float x000, float x100, float x010, float x110, float x001, float x101, float x011, float x111) {
    const float p = lerp(a2, lerp(a1, x000, x100), lerp(a1, x010, x110));
    const float q = lerp(a2, lerp(a1, x001, x101), lerp(a1, x011, x111));
    return lerp(a3, p, q);
}
float smoothstep(float x) {   // x * x * x * (x * (x * 6 - 15) + 10)
    float inner = x * 6.0f;
    inner = inner - 15.0f;
    inner = x * inner;
    inner = inner + 10.0f;
    const float cube = x * x * x;
    return cube * inner;
}

// ---- hashes ----------------------------------------------------------------------------------------
namespace {
uint32_t rol(uint32_t v, int k) { return (v << k) | (v >> (32 - k)); }
uint32_t ror(uint32_t v, int k) { return (v >> k) | (v << (32 - k)); }
}  // namespace

void md5(const std::string& s, uint8_t out[16]) {
    // RFC 1321. Per-round sine table and shift amounts.
    uint32_t K[64];
    for (int i = 0; i < 64; ++i) K[i] = static_cast<uint32_t>(std::floor(std::fabs(std::sin(static_cast<double>(i + 1))) * 4294967296.0));
    const int S[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23}, {6, 10, 15, 21}};
    std::vector<uint8_t> m(s.begin(), s.end());
    const uint64_t bits = static_cast<uint64_t>(m.size()) * 8;
    m.push_back(0x80);
    while (m.size() % 64 != 56) m.push_back(0);
    for (int i = 0; i < 8; ++i) m.push_back(static_cast<uint8_t>(bits >> (8 * i)));
    uint32_t h[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
    for (size_t off = 0; off < m.size(); off += 64) {
        uint32_t w[16];
        for (int i = 0; i < 16; ++i) std::memcpy(&w[i], &m[off + 4 * i], 4);   // little-endian words (x86 host)
        uint32_t a = h[0], b = h[1], c = h[2], d = h[3];
        for (int i = 0; i < 64; ++i) {
            const int round = i / 16;
            uint32_t f;
            int g;
            switch (round) {
                case 0: f = (b & c) | (~b & d); g = i; break;
                case 1: f = (d & b) | (~d & c); g = (5 * i + 1) & 15; break;
                case 2: f = b ^ c ^ d; g = (3 * i + 5) & 15; break;
                default: f = c ^ (b | ~d); g = (7 * i) & 15; break;
            }
            const uint32_t nb = b + rol(a + f + K[i] + w[g], S[round][i & 3]);
            a = d;
            d = c;
            c = b;
            b = nb;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d;
    }
    std::memcpy(out, h, 16);
}

void sha256(const uint8_t* msg, size_t len, uint8_t out[32]) {
    // FIPS 180-4. Round constants: fractional parts of the cube roots of the first 64 primes.
    static uint32_t K[64];
    static uint32_t H0[8];
    static bool init = false;
    if (!init) {
        int primes[64], np = 0;
        for (int c = 2; np < 64; ++c) {
            bool p = true;
            for (int d = 2; d * d <= c; ++d) p &= (c % d) != 0;
            if (p) primes[np++] = c;
        }
        for (int i = 0; i < 64; ++i) {
            const long double r = std::cbrt(static_cast<long double>(primes[i]));
            K[i] = static_cast<uint32_t>((r - std::floor(r)) * 4294967296.0L);
        }
        for (int i = 0; i < 8; ++i) {
            const long double r = std::sqrt(static_cast<long double>(primes[i]));
            H0[i] = static_cast<uint32_t>((r - std::floor(r)) * 4294967296.0L);
        }
        init = true;
    }
    std::vector<uint8_t> m(msg, msg + len);
    const uint64_t bits = static_cast<uint64_t>(len) * 8;
    m.push_back(0x80);
    while (m.size() % 64 != 56) m.push_back(0);
    for (int i = 7; i >= 0; --i) m.push_back(static_cast<uint8_t>(bits >> (8 * i)));
    uint32_t h[8];
    std::memcpy(h, H0, sizeof h);
    for (size_t off = 0; off < m.size(); off += 64) {
        uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = (uint32_t{m[off + 4 * i]} << 24) | (uint32_t{m[off + 4 * i + 1]} << 16) | (uint32_t{m[off + 4 * i + 2]} << 8) | m[off + 4 * i + 3];
        for (int i = 16; i < 64; ++i)
            w[i] = w[i - 16] + (ror(w[i - 15], 7) ^ ror(w[i - 15], 18) ^ (w[i - 15] >> 3)) + w[i - 7] + (ror(w[i - 2], 17) ^ ror(w[i - 2], 19) ^ (w[i - 2] >> 10));
        uint32_t v[8];
        std::memcpy(v, h, sizeof v);
        for (int i = 0; i < 64; ++i) {
            const uint32_t t1 = v[7] + (ror(v[4], 6) ^ ror(v[4], 11) ^ ror(v[4], 25)) + ((v[4] & v[5]) ^ (~v[4] & v[6])) + K[i] + w[i];
            const uint32_t t2 = (ror(v[0], 2) ^ ror(v[0], 13) ^ ror(v[0], 22)) + ((v[0] & v[1]) ^ (v[0] & v[2]) ^ (v[1] & v[2]));
            for (int k = 7; k > 0; --k) v[k] = v[k - 1];
            v[4] += t1;
            v[0] = t1 + t2;
        }
        for (int i = 0; i < 8; ++i) h[i] += v[i];
    }
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 4; ++j) out[4 * i + j] = static_cast<uint8_t>(h[i] >> (24 - 8 * j));
}

int64_t zoom_seed_of(int64_t world_seed) {
    // Guava Hashing.sha256().hashLong(seed).asLong(): little-endian in, first 8 bytes LE out.
    uint8_t in[8], d[32];
    for (int i = 0; i < 8; ++i) in[i] = static_cast<uint8_t>(static_cast<uint64_t>(world_seed) >> (8 * i));
    sha256(in, 8, d);
    uint64_t r = 0;
    for (int i = 0; i < 8; ++i) r |= uint64_t{d[i]} << (8 * i);
    return static_cast<int64_t>(r);
}

// ---- Xoroshiro128++ sources ---------------------------------------------------------------------------
uint64_t mix13(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

Xrand::Xrand(uint64_t lo, uint64_t hi) : s0(lo), s1(hi) {
    if (s0 == 0 && s1 == 0) {
        s0 = 0x9E3779B97F4A7C15ull;   // golden ratio
        s1 = 0x6A09E667F3BCC909ull;   // silver ratio
    }
}

Xrand Xrand::from_seed(int64_t seed) {
    const uint64_t lo = static_cast<uint64_t>(seed) ^ 0x6A09E667F3BCC909ull;
    return Xrand(mix13(lo), mix13(lo + 0x9E3779B97F4A7C15ull));
}

uint64_t Xrand::next() {
    const uint64_t a = s0, b = s1;
    const uint64_t sum = a + b;
    const uint64_t out = ((sum << 17) | (sum >> 47)) + a;
    const uint64_t x = b ^ a;
    s0 = ((a << 49) | (a >> 15)) ^ x ^ (x << 21);
    s1 = (x << 28) | (x >> 36);
    return out;
}

int32_t Xrand::next_int(int32_t bound) {
    // Lemire: multiply a 32-bit draw by the bound; reject low parts below 2^32 mod bound
    const uint64_t ub = static_cast<uint32_t>(bound);
    uint64_t m = (next() & 0xFFFFFFFFull) * ub;
    if ((m & 0xFFFFFFFFull) < ub) {
        const uint64_t threshold = ((1ull << 32) - ub) % ub;
        while ((m & 0xFFFFFFFFull) < threshold) m = (next() & 0xFFFFFFFFull) * ub;
    }
    return static_cast<int32_t>(m >> 32);
}

double Xrand::next_double() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
float Xrand::next_float() { return static_cast<float>(next() >> 40) * 0x1.0p-24f; }

Factory Factory::fork(Xrand& r) {
    Factory f;
    f.lo = r.next();
    f.hi = r.next();
    return f;
}

Xrand Factory::hash(const std::string& name) const {
    uint8_t d[16];
    md5(name, d);
    uint64_t a = 0, b = 0;
    for (int i = 0; i < 8; ++i) a = (a << 8) | d[i];
    for (int i = 8; i < 16; ++i) b = (b << 8) | d[i];
    return Xrand(a ^ lo, b ^ hi);
}

int64_t block_seed(int32_t x, int32_t y, int32_t z) {
    uint64_t s = static_cast<uint64_t>(static_cast<int64_t>(wmul(x, 3129871)));
    s ^= static_cast<uint64_t>(static_cast<int64_t>(z)) * 116129781ull;
    s ^= static_cast<uint64_t>(static_cast<int64_t>(y));
    s = s * s * 42317861ull + s * 11ull;
    return static_cast<int64_t>(s) >> 16;
}

Xrand Factory::at(int32_t x, int32_t y, int32_t z) const { return Xrand(static_cast<uint64_t>(block_seed(x, y, z)) ^ lo, hi); }

// ---- java.util.Random -------------------------------------------------------------------------------
int32_t Lcg::next(int bits) {
    s = (s * 0x5DEECE66Dull + 0xBull) & ((1ull << 48) - 1);
    return static_cast<int32_t>(static_cast<uint32_t>(s >> (48 - bits)));
}
int32_t Lcg::next_int(int32_t bound) {
    if ((bound & -bound) == bound) return static_cast<int32_t>((static_cast<int64_t>(bound) * next(31)) >> 31);
    for (;;) {
        const int32_t bits = next(31);
        const int32_t val = bits % bound;
        if (static_cast<int32_t>(static_cast<uint32_t>(bits) - static_cast<uint32_t>(val) + static_cast<uint32_t>(bound - 1)) >= 0) return val;
    }
}
double Lcg::next_double() {
    const int64_t hi = next(26), lo = next(27);
    return static_cast<double>((hi << 27) + lo) * 0x1.0p-53;
}

// ---- gradient noise ------------------------------------------------------------------------------------
namespace {
// SimplexNoise.GRADIENT
const int G[16][3] = {{1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1},
                      {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}, {1, 1, 0}, {0, -1, 1}, {-1, 1, 0}, {0, -1, -1}};
inline int hp(const Octave& o, int i) { return o.perm[i & 255]; }

struct Corners {
    int c[8];   // index bits: x << 2 | y << 1 | z
};
Corners lattice(const Octave& o, int x, int y, int z) {
    Corners k;
    for (int dx = 0; dx < 2; ++dx)
        for (int dy = 0; dy < 2; ++dy) {
            const int xy = hp(o, wadd(hp(o, wadd(x, dx)), wadd(y, dy)));
            for (int dz = 0; dz < 2; ++dz) k.c[(dx << 2) | (dy << 1) | dz] = hp(o, wadd(xy, wadd(z, dz)));
        }
    return k;
}

// point path: (gx*x + gy*y) + gz*z, then lerp3 with smoothstep weights (y weight from ty)
float point_sample(const Octave& o, int fx, int fy, int fz, float rx, float ry, float rz, float ty) {
    const Corners k = lattice(o, fx, fy, fz);
    float d[8];
    for (int i = 0; i < 8; ++i) {
        const float px = (i & 4) ? rx - 1.0f : rx, py = (i & 2) ? ry - 1.0f : ry, pz = (i & 1) ? rz - 1.0f : rz;
        const int* g = G[k.c[i] & 15];
        const float ab = static_cast<float>(g[0]) * px + static_cast<float>(g[1]) * py;
        d[i] = ab + static_cast<float>(g[2]) * pz;
    }
    return lerp3(smoothstep(rx), smoothstep(ty), smoothstep(rz), d[0], d[4], d[2], d[6], d[1], d[5], d[3], d[7]);
}

// volume path: (gx*x + gz*z) + gy*y
float volume_sample(const Octave& o, int fx, int fy, int fz, float rx, float ry, float rz, float ax, float ay, float az) {
    const Corners k = lattice(o, fx, fy, fz);
    float d[8];
    for (int i = 0; i < 8; ++i) {
        const float px = (i & 4) ? rx - 1.0f : rx, py = (i & 2) ? ry - 1.0f : ry, pz = (i & 1) ? rz - 1.0f : rz;
        const int* g = G[k.c[i] & 15];
        const float xz = static_cast<float>(g[0]) * px + static_cast<float>(g[2]) * pz;
        d[i] = xz + static_cast<float>(g[1]) * py;
    }
    return lerp3(ax, ay, az, d[0], d[4], d[2], d[6], d[1], d[5], d[3], d[7]);
}

double fudge(double orig_y, double rel_y, double smear) {   // SmearedPerlinNoise.computeFudgeY
    const double lim = (orig_y >= 0.0 && orig_y < rel_y) ? orig_y : rel_y;
    return static_cast<double>(floor_i(lim / smear + static_cast<double>(1.0E-7f))) * smear;
}
}  // namespace

double wrap_coord(double x) {
    const double limit = 0x1.fffffffffffffp+23;   // Math.nextDown(2^24)
    if (x >= -limit && x < limit) return x;
    return x - std::floor(x / 33554432.0 + 0.5) * 33554432.0;
}

float noise_get(const Octave& o, double x, double y, double z) {
    const double X = wrap_coord(x) + o.xo, Y = wrap_coord(y) + o.yo, Z = wrap_coord(z) + o.zo;
    const int fx = floor_i(X), fy = floor_i(Y), fz = floor_i(Z);
    const float rx = static_cast<float>(X - fx), ry = static_cast<float>(Y - fy), rz = static_cast<float>(Z - fz);
    return point_sample(o, fx, fy, fz, rx, ry, rz, ry);
}

float noise_get_smeared(const Octave& o, double x, double y, double z) {
    const double X = wrap_coord(x) + o.xo, Y = wrap_coord(y) + o.yo, Z = wrap_coord(z) + o.zo;
    const int fx = floor_i(X), fy = floor_i(Y), fz = floor_i(Z);
    const double rel = Y - fy;
    const float rx = static_cast<float>(X - fx), rz = static_cast<float>(Z - fz);
    return point_sample(o, fx, fy, fz, rx, static_cast<float>(rel - fudge(y, rel, o.smear)), rz, static_cast<float>(rel));
}

float noise_add_volume(const Octave& o, int bx, int by, int bz, double xz, double ys) {
    const double X = wrap_coord(bx * xz) + o.xo, Y = wrap_coord(by * ys) + o.yo, Z = wrap_coord(bz * xz) + o.zo;
    const int fx = floor_i(X), fy = floor_i(Y), fz = floor_i(Z);
    const float rx = static_cast<float>(X - fx), ry = static_cast<float>(Y - fy), rz = static_cast<float>(Z - fz);
    return volume_sample(o, fx, fy, fz, rx, ry, rz, smoothstep(rx), smoothstep(ry), smoothstep(rz));
}

float noise_add_volume_smeared(const Octave& o, int bx, int by, int bz, double xz, double ys) {
    const double orig = by * ys;
    const double X = wrap_coord(bx * xz) + o.xo, Y = wrap_coord(orig) + o.yo, Z = wrap_coord(bz * xz) + o.zo;
    const int fx = floor_i(X), fy = floor_i(Y), fz = floor_i(Z);
    const double rel = Y - fy;
    const float rx = static_cast<float>(X - fx), rz = static_cast<float>(Z - fz);
    const float ry = static_cast<float>(rel - fudge(orig, rel, o.smear));
    return volume_sample(o, fx, fy, fz, rx, ry, rz, smoothstep(rx), smoothstep(static_cast<float>(rel)), smoothstep(rz));
}

// ---- JSON -----------------------------------------------------------------------------------------------
namespace {
struct Reader {
    const std::string& s;
    size_t i = 0;
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
    }
    Json value() {
        ws();
        if (i >= s.size()) throw std::runtime_error("json: unexpected end");
        Json j;
        const char c = s[i];
        if (c == '{') {
            j.kind = Json::Obj;
            ++i;
            ws();
            if (s[i] == '}') { ++i; return j; }
            for (;;) {
                ws();
                Json k = value();
                ws();
                if (s[i++] != ':') throw std::runtime_error("json: expected ':'");
                Json v = value();
                j.o.emplace_back(k.s, std::move(v));
                ws();
                if (s[i] == ',') { ++i; continue; }
                if (s[i] == '}') { ++i; return j; }
                throw std::runtime_error("json: bad object");
            }
        }
        if (c == '[') {
            j.kind = Json::Arr;
            ++i;
            ws();
            if (s[i] == ']') { ++i; return j; }
            for (;;) {
                j.a.push_back(value());
                ws();
                if (s[i] == ',') { ++i; continue; }
                if (s[i] == ']') { ++i; return j; }
                throw std::runtime_error("json: bad array");
            }
        }
        if (c == '"') {
            j.kind = Json::Str;
            ++i;
            while (s[i] != '"') {
                if (s[i] == '\\') {
                    ++i;
                    const char e = s[i++];
                    if (e == 'n') j.s += '\n';
                    else if (e == 't') j.s += '\t';
                    else if (e == 'u') { j.s += '?'; i += 4; }
                    else j.s += e;
                } else {
                    j.s += s[i++];
                }
            }
            ++i;
            return j;
        }
        if (s.compare(i, 4, "true") == 0) { j.kind = Json::Bool; j.b = true; i += 4; return j; }
        if (s.compare(i, 5, "false") == 0) { j.kind = Json::Bool; i += 5; return j; }
        if (s.compare(i, 4, "null") == 0) { i += 4; return j; }
        j.kind = Json::Num;
        const size_t b = i;
        while (i < s.size() && (std::isdigit(static_cast<unsigned char>(s[i])) || s[i] == '-' || s[i] == '+' || s[i] == '.' || s[i] == 'e' || s[i] == 'E')) ++i;
        j.s = s.substr(b, i - b);
        return j;
    }
};
}  // namespace

const Json& Json::operator[](const std::string& k) const {
    for (const auto& kv : o)
        if (kv.first == k) return kv.second;
    throw std::runtime_error("json: missing key " + k);
}
bool Json::has(const std::string& k) const {
    for (const auto& kv : o)
        if (kv.first == k) return true;
    return false;
}
int64_t Json::i64() const {
    if (kind == Bool) return b ? 1 : 0;
    return std::stoll(s);
}
double Json::hexd() const { return bd(std::stoull(s, nullptr, 16)); }

Json parse_json_file(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    Reader r{text};
    return r.value();
}

}  // namespace orc
