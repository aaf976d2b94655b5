// Oracle: per-seed state, chunk biomes, NoiseBasedAquifer fill and the material-rule surface
// pass. Restates cpu/src/terrain.cpp:14-812 (and cpu/src/biome.cpp:272-289 for the zoomed
// biome lookup) in plain single-threaded C++.
#include <algorithm>
#include <climits>
#include <functional>
#include <stdexcept>

#include "oracle.h"

namespace orc {

namespace {

constexpr int kWayBelow = -2032 * 16;   // DimensionType.WAY_BELOW_MIN_Y

// SimplexNoise over (x, y) with z = 0; offsets are nextDouble() * 0.0 = 0 (terrain.cpp:14-70)
const int8_t kSimplexGrad[12][3] = {{1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1},
                                    {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};

void simplex_perm(Lcg& r, uint8_t* p) {
    for (int k = 0; k < 3; ++k) (void)r.next_double();
    for (int i = 0; i < 256; ++i) p[i] = static_cast<uint8_t>(i);
    for (int i = 0; i < 256; ++i) std::swap(p[i], p[i + r.next_int(256 - i)]);
}

double simplex_term(int g, double x, double y) {
    double t = 0.5 - x * x - y * y - 0.0 * 0.0;
    if (t < 0.0) return 0.0;
    t *= t;
    const double dot = (static_cast<double>(kSimplexGrad[g][0]) * x + static_cast<double>(kSimplexGrad[g][1]) * y) + static_cast<double>(kSimplexGrad[g][2]) * 0.0;
    return t * t * dot;
}

float simplex(const uint8_t* p, double px, double py) {
    const double x = px + 0.0, y = py + 0.0;   // the zero offsets of GradientNoise(random, 0.0)
    const double s3 = std::sqrt(3.0), F2 = 0.5 * (s3 - 1.0), G2 = (3.0 - s3) / 6.0;
    const double s = (x + y) * F2;
    const int i = floor_i(x + s), j = floor_i(y + s);
    const double t = static_cast<double>(wadd(i, j)) * G2;
    const double x0 = x - (static_cast<double>(i) - t), y0 = y - (static_cast<double>(j) - t);
    const int i1 = x0 > y0 ? 1 : 0, j1 = 1 - i1;
    const double x1 = x0 - i1 + G2, y1 = y0 - j1 + G2, x2 = x0 - 1.0 + 2.0 * G2, y2 = y0 - 1.0 + 2.0 * G2;
    auto P = [&](int v) { return static_cast<int>(p[v & 0xFF]); };
    const int ii = i & 0xFF, jj = j & 0xFF;
    const double n = simplex_term(P(ii + P(jj)) % 12, x0, y0) + simplex_term(P(ii + i1 + P(jj + j1)) % 12, x1, y1) +
                     simplex_term(P(ii + 1 + P(jj + 1)) % 12, x2, y2);
    return static_cast<float>(70.0 * n);
}

struct Fluid {
    int level, type;
    int at(int y, int air) const { return y < level ? type : air; }
};

// NoiseBasedAquifer over one chunk volume (terrain.cpp:200-435)
class Aquifer {
public:
    Aquifer(World& w, Density& d, const Vol& v) : w_(w), d_(d), m_(w.m) {
        const auto& roots = w.ir.roots;
        barrier_ = roots.at("aquifer/barrier");
        flood_ = roots.at("aquifer/fluid_level_floodedness");
        spread_ = roots.at("aquifer/fluid_level_spread");
        lava_ = roots.at("aquifer/lava");
        excl_ = roots.at("aquifer/exclusion");
        surf_ = roots.at("aquifer/surface_level");
        const int max_x = v.x0 + (v.sx - 1) * v.dx, max_y = v.y0 + (v.sy - 1) * v.dy, max_z = v.z0 + (v.sz - 1) * v.dz;
        gx0_ = (v.x0 - 5) >> 4;
        const int gx1 = ((max_x - 5) >> 4) + 1;
        nx_ = gx1 - gx0_ + 1;
        gy0_ = floor_div(v.y0 + 1, 12) - 1;
        const int gy1 = floor_div(max_y + 1, 12) + 1;
        gz0_ = (v.z0 - 5) >> 4;
        const int gz1 = ((max_z - 5) >> 4) + 1;
        nz_ = gz1 - gz0_ + 1;
        const size_t cells = static_cast<size_t>(nx_) * static_cast<size_t>(gy1 - gy0_ + 1) * static_cast<size_t>(nz_);
        loc_.assign(cells, INT64_MAX);
        status_.assign(cells, Fluid{0, -1});
        known_.assign(cells, 0);
        // the highest surface level over the quart grid of the window, sampled as one volume
        const int qx0 = (gx0_ * 16) >> 2, qx1 = (gx1 * 16 + 9) >> 2, qz0 = (gz0_ * 16) >> 2, qz1 = (gz1 * 16 + 9) >> 2;
        const Vol sv{qx1 - qx0 + 1, 1, qz1 - qz0 + 1, qx0 * 4, 0, qz0 * 4, 4, 1, 4};
        std::vector<float> b(static_cast<size_t>(sv.n()));
        d_.sample(surf_, sv, b.data());
        int top = INT_MIN;
        for (int z = 0; z < sv.sz; ++z)
            for (int x = 0; x < sv.sx; ++x) {
                const int level = floor_i(static_cast<double>(b[static_cast<size_t>(sv.at(x, 0, z))]));
                surface_[key(sv.x0 + x * 4, sv.z0 + z * 4)] = level;
                top = std::max(top, level);
            }
        skip_above_ = (floor_div(top + 8 + 12, 12) + 1) * 12 + 11 - 1;
    }

    // computeSubstance: the state, or -1 where the default block stays
    int substance(int x, int y, int z, double density) {
        if (density > 0.0) return -1;
        const Fluid global = global_fluid(y);
        if (y > skip_above_) return global.at(y, m_.air);
        if (global.at(y, m_.air) == m_.lava) return m_.lava;
        const int ax = (x - 5) >> 4, ay = floor_div(y + 1, 12), az = (z - 5) >> 4;
        int dist[4] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX}, cell[4] = {0, 0, 0, 0};
        for (int ox = 0; ox <= 1; ++ox)
            for (int oy = -1; oy <= 1; ++oy)
                for (int oz = 0; oz <= 1; ++oz) {
                    const int gx = ax + ox, gy = ay + oy, gz = az + oz;
                    const int c = index(gx, gy, gz);
                    int lx, ly, lz;
                    location(c, gx, gy, gz, lx, ly, lz);
                    const int dd = (lx - x) * (lx - x) + (ly - y) * (ly - y) + (lz - z) * (lz - z);
                    // keep the four nearest; a later cell at an equal distance goes first
                    int k = 0;
                    while (k < 4 && dist[k] < dd) ++k;
                    if (k == 4) continue;
                    for (int j = 3; j > k; --j) { dist[j] = dist[j - 1]; cell[j] = cell[j - 1]; }
                    dist[k] = dd;
                    cell[k] = c;
                }
        const Fluid s1 = status(cell[0]);
        const double sim12 = similarity(dist[0], dist[1]);
        const int fluid = s1.at(y, m_.air);
        if (sim12 <= 0.0) return fluid;
        if (fluid == m_.water && global_fluid(y - 1).at(y - 1, m_.air) == m_.lava) return fluid;
        bool have_barrier = false;
        double barrier = 0.0;
        const Fluid s2 = status(cell[1]);
        if (density + sim12 * pressure(x, y, z, have_barrier, barrier, s1, s2) > 0.0) return -1;
        const Fluid s3 = status(cell[2]);
        const double sim13 = similarity(dist[0], dist[2]);
        if (sim13 > 0.0 && density + sim12 * sim13 * pressure(x, y, z, have_barrier, barrier, s1, s3) > 0.0) return -1;
        const double sim23 = similarity(dist[1], dist[2]);
        if (sim23 > 0.0 && density + sim12 * sim23 * pressure(x, y, z, have_barrier, barrier, s2, s3) > 0.0) return -1;
        return fluid;
    }

private:
    static int64_t key(int x, int z) { return (static_cast<int64_t>(z) << 32) | static_cast<uint32_t>(x); }
    static double similarity(int a, int b) { return 1.0 - static_cast<double>(b - a) / 25.0; }

    Fluid global_fluid(int y) const {
        return y < std::min(-54, m_.sea) ? Fluid{-54, m_.lava} : Fluid{m_.sea, m_.default_fluid};
    }
    int index(int gx, int gy, int gz) const { return ((gy - gy0_) * nz_ + (gz - gz0_)) * nx_ + (gx - gx0_); }

    // the cell's random location, stored as BlockPos.asLong (26/12/26 bits) like vanilla
    void location(int c, int gx, int gy, int gz, int& x, int& y, int& z) {
        int64_t& p = loc_[static_cast<size_t>(c)];
        if (p == INT64_MAX) {
            Xrand r = w_.aquifer.at(gx, gy, gz);
            const int px = gx * 16 + r.next_int(10);
            const int py = gy * 12 + r.next_int(9);
            const int pz = gz * 16 + r.next_int(10);
            p = static_cast<int64_t>(((static_cast<uint64_t>(static_cast<int64_t>(px)) & 0x3FFFFFFull) << 38) |
                                     ((static_cast<uint64_t>(static_cast<int64_t>(pz)) & 0x3FFFFFFull) << 12) |
                                     (static_cast<uint64_t>(static_cast<int64_t>(py)) & 0xFFFull));
        }
        x = static_cast<int>(p >> 38);
        y = static_cast<int>((p << 52) >> 52);
        z = static_cast<int>((p << 26) >> 38);
    }

    int surface_level(int bx, int bz) {
        const int qx = (bx >> 2) << 2, qz = (bz >> 2) << 2;
        auto it = surface_.find(key(qx, qz));
        if (it != surface_.end()) return it->second;
        const int level = floor_i(static_cast<double>(d_.point(surf_, qx, 0, qz)));
        surface_[key(qx, qz)] = level;
        return level;
    }

    Fluid status(int c) {
        if (known_[static_cast<size_t>(c)]) return status_[static_cast<size_t>(c)];
        const int64_t p = loc_[static_cast<size_t>(c)];
        const int x = static_cast<int>(p >> 38), y = static_cast<int>((p << 52) >> 52), z = static_cast<int>((p << 26) >> 38);
        const Fluid s = compute_fluid(x, y, z);
        status_[static_cast<size_t>(c)] = s;
        known_[static_cast<size_t>(c)] = 1;
        return s;
    }

    Fluid compute_fluid(int x, int y, int z) {
        static const int off[13][2] = {{0, 0}, {-2, -1}, {-1, -1}, {0, -1}, {1, -1}, {-3, 0}, {-2, 0}, {-1, 0}, {1, 0}, {-2, 1}, {-1, 1}, {0, 1}, {1, 1}};
        const Fluid global = global_fluid(y);
        int lowest = INT_MAX;
        bool centre_under = false;
        for (const auto& o : off) {
            const int sx = x + o[0] * 16, sz = z + o[1] * 16;
            const int level = surface_level(sx, sz), adj = level + 8;
            const bool start = o[0] == 0 && o[1] == 0;
            if (start && y - 12 > adj) return global;
            const bool pokes = y + 12 > adj;
            if (pokes || start) {
                const Fluid at = global_fluid(adj);
                if (at.at(adj, m_.air) != m_.air) {
                    if (start) centre_under = true;
                    if (pokes) return at;
                }
            }
            lowest = std::min(lowest, level);
        }
        const int level = fluid_level(x, y, z, global, lowest, centre_under);
        int type = global.type;
        if (level <= -10 && level != kWayBelow && global.type != m_.lava) {
            const double lv = static_cast<double>(d_.point(lava_, floor_div(x, 64), floor_div(y, 40), floor_div(z, 64)));
            if (std::fabs(lv) > 0.3) type = m_.lava;
        }
        return Fluid{level, type};
    }

    int fluid_level(int x, int y, int z, const Fluid& global, int lowest, bool centre_under) {
        double partially = -1.0, fully = -1.0;
        if (!(d_.point(excl_, x, y, z) > 0.0f)) {
            double factor = 0.0;
            if (centre_under) {   // Mth.clampedMap(below, 0, 64, 1, 0)
                const double f = static_cast<double>(lowest + 8 - y) / 64.0;
                factor = f < 0.0 ? 1.0 : (f > 1.0 ? 0.0 : 1.0 + f * (0.0 - 1.0));
            }
            const double n = dclamp(static_cast<double>(d_.point(flood_, x, y, z)), -1.0, 1.0);
            const double u = (factor - 1.0) / (0.0 - 1.0);
            fully = n - (-0.3 + u * (0.8 - -0.3));
            partially = n - (-0.8 + u * (0.4 - -0.8));
        }
        if (fully > 0.0) return global.level;
        if (partially > 0.0) {
            const int cy = floor_div(y, 40);
            const double s = static_cast<double>(d_.point(spread_, floor_div(x, 16), cy, floor_div(z, 16)) * 10.0f);
            return std::min(lowest, cy * 40 + 20 + floor_i(s / 3.0) * 3);
        }
        return kWayBelow;
    }

    double pressure(int x, int y, int z, bool& have, double& barrier, const Fluid& a, const Fluid& b) {
        const int ta = a.at(y, m_.air), tb = b.at(y, m_.air);
        if ((ta == m_.lava && tb == m_.water) || (ta == m_.water && tb == m_.lava)) return 2.0;
        const int diff = std::abs(a.level - b.level);
        if (diff == 0) return 0.0;
        const double above = static_cast<double>(y) + 0.5 - 0.5 * static_cast<double>(a.level + b.level);
        const double towards = static_cast<double>(diff) / 2.0 - std::fabs(above);
        double g;
        if (above > 0.0) {
            const double c = 0.0 + towards;
            g = c > 0.0 ? c / 1.5 : c / 2.5;
        } else {
            const double c = 3.0 + towards;
            g = c > 0.0 ? c / 3.0 : c / 10.0;
        }
        double n = 0.0;
        if (!(g < -2.0) && !(g > 2.0)) {
            if (!have) { barrier = static_cast<double>(d_.point(barrier_, x, y, z)); have = true; }
            n = barrier;
        }
        return 2.0 * (n + g);
    }

    World& w_;
    Density& d_;
    const Model& m_;
    int barrier_, flood_, spread_, lava_, excl_, surf_;
    int gx0_, gy0_, gz0_, nx_, nz_, skip_above_;
    std::vector<int64_t> loc_;
    std::vector<Fluid> status_;
    std::vector<uint8_t> known_;
    std::unordered_map<int64_t, int> surface_;
};

// SurfaceSystem.buildSurface + the material rules of one chunk (terrain.cpp:437-759)
class Surface {
public:
    using QuartBiome = std::function<int(int, int, int)>;
    Surface(World& w, Density& d, int cx, int cz, uint16_t* blocks, int* hm, const QuartBiome& qb)
        : w_(w), d_(d), m_(w.m), cx_(cx), cz_(cz), blocks_(blocks), hm_(hm), qb_(qb) {
        ymin_ = w.ir.min_y;
        ymax_ = w.ir.min_y + w.ir.height - 1;
        const Vol pv{16, 1, 16, cx * 16, 0, cz * 16, 1, 1, 1};
        prelim_.resize(256);
        d_.sample(w.ir.roots.at("surface/chunk_surface_level"), pv, prelim_.data());
    }

    void run() {
        for (int x = 0; x < 16; ++x)
            for (int z = 0; z < 16; ++z) column(x, z);
    }

private:
    double nz(const char* name, double x, double y, double z) const { return static_cast<double>(stack_get(w_.tables, w_.tables.noise.at(name), x, y, z)); }
    int H(int x, int z) const { return hm_[x * 16 + z] - 1; }
    size_t at(int y) const { return static_cast<size_t>(((y - ymin_) * 16 + lz_) * 16 + lx_); }
    int get(int y) const { return (y < ymin_ || y > ymax_) ? m_.air : blocks_[at(y)]; }
    bool air(int s) const { return m_.is_air[static_cast<size_t>(s)] != 0; }
    bool fluid(int s) const { return m_.is_water[static_cast<size_t>(s)] || m_.is_lava[static_cast<size_t>(s)]; }

    void set(int y, int s) {   // setBlockState + Heightmap.update(WORLD_SURFACE_WG)
        if (y < ymin_ || y > ymax_) return;
        blocks_[at(y)] = static_cast<uint16_t>(s);
        int& h = hm_[lx_ * 16 + lz_];
        const int first = h;
        if (y <= first - 2) return;
        if (!air(s)) {
            if (y >= first) h = y + 1;
        } else if (first - 1 == y) {
            for (int yy = y - 1; yy >= ymin_; --yy)
                if (!air(blocks_[at(yy)])) { h = yy + 1; return; }
            h = ymin_;
        }
    }

    int biome_at(int x, int y, int z) const {
        int qx, qy, qz;
        zoom_quart(w_.zseed, x, y, z, qx, qy, qz);
        return qb_(qx, qy, qz);
    }

    void column(int x, int z) {
        lx_ = x;
        lz_ = z;
        const int bx = cx_ * 16 + x, bz = cz_ * 16 + z;
        const int start = H(x, z) + 1;
        const int top_biome = biome_at(bx, start, bz);
        if (top_biome == m_.eroded_badlands) eroded_badlands(bx, bz, start);
        const int height = H(x, z) + 1;
        bx_ = bx;
        bz_ = bz;
        gx_ = H(std::min(x + 1, 15), z) - H(std::max(x - 1, 0), z);
        gz_ = H(x, std::min(z + 1, 15)) - H(x, std::max(z - 1, 0));
        depth_ = to_int(nz("minecraft:surface", bx, 0.0, bz) * 2.75 + 3.0 + w_.random.at(bx, 0, bz).next_double() * 0.25);
        have_secondary_ = have_min_surface_ = false;
        int stone_above = 0, water = INT_MIN, ceiling = INT_MAX;
        for (int y = height; y >= ymin_; --y) {
            const int old = get(y);
            if (air(old)) {
                stone_above = 0;
                water = INT_MIN;
            } else if (fluid(old)) {
                if (water == INT_MIN) water = y + 1;
            } else {
                if (ceiling >= y) {
                    ceiling = kWayBelow;
                    for (int k = y - 1; k >= ymin_ - 1; --k) {
                        const int s = get(k);
                        if (air(s) || fluid(s)) { ceiling = k + 1; break; }
                    }
                }
                ++stone_above;
                above_ = stone_above;
                below_ = y - ceiling + 1;
                water_ = water;
                by_ = y;
                have_biome_ = false;
                if (y >= ymin_ && y <= ymax_) {
                    const int s = rule(w_.m.root);
                    if (s >= 0) set(y, s);
                }
            }
        }
        if (top_biome == m_.frozen_ocean || top_biome == m_.deep_frozen_ocean) frozen_ocean(min_surface(), top_biome, bx, bz, start);
    }

    double secondary() {
        if (!have_secondary_) { secondary_ = nz("minecraft:surface_secondary", bx_, 0.0, bz_); have_secondary_ = true; }
        return secondary_;
    }
    int min_surface() {
        if (!have_min_surface_) {
            min_surface_ = floor_i(static_cast<double>(prelim_[static_cast<size_t>(lz_ * 16 + lx_)])) + depth_ - 8;
            have_min_surface_ = true;
        }
        return min_surface_;
    }
    int biome() {
        if (!have_biome_) { biome_ = biome_at(bx_, by_, bz_); have_biome_ = true; }
        return biome_;
    }
    float temperature(int b, int x, int y, int z) const { return biome_temperature(w_, b, x, y, z); }

    bool cond(int ci) {
        const Cond& c = m_.conds[static_cast<size_t>(ci)];
        const std::string& t = c.t;
        if (t == "biome") { const int b = biome(); return ((c.mask[b >> 6] >> (b & 63)) & 1ull) != 0; }
        if (t == "noise_threshold") {
            const double v = c.is3d ? nz(c.noise.c_str(), bx_, by_, bz_) : nz(c.noise.c_str(), bx_, 0.0, bz_);
            return v >= c.mn && v <= c.mx;
        }
        if (t == "vertical_gradient") {
            if (by_ <= c.true_at) return true;
            if (by_ >= c.false_at) return false;
            const double u = (static_cast<double>(by_) - c.true_at) / static_cast<double>(c.false_at - c.true_at);
            Xrand r = w_.gradient.at(c.random_name).at(bx_, by_, bz_);
            return static_cast<double>(r.next_float()) < 1.0 + u * (0.0 - 1.0);
        }
        if (t == "y_above") return by_ + (c.add_stone ? above_ : 0) >= c.anchor + depth_ * c.mult;
        if (t == "water") return water_ == INT_MIN || by_ + (c.add_stone ? above_ : 0) >= water_ + c.offset + depth_ * c.mult;
        if (t == "temperature") return !(temperature(biome(), bx_, by_, bz_) >= 0.15f);
        if (t == "steep") return gx_ <= -4 || gz_ >= 4;
        if (t == "not") return !cond(c.child);
        if (t == "hole") return depth_ <= 0;
        if (t == "above_preliminary_surface") return by_ >= min_surface();
        if (t == "stone_depth") {
            int extra = 0;
            if (c.secondary != 0) extra = to_int(0.0 + (secondary() - -1.0) / (1.0 - -1.0) * (static_cast<double>(c.secondary) - 0.0));
            return (c.ceiling ? below_ : above_) <= 1 + c.offset + (c.add_surface ? depth_ : 0) + extra;
        }
        throw std::runtime_error("condition " + t);
    }

    float ore_value(const std::string& root) {
        auto it = ore_.find(root);
        if (it == ore_.end()) {
            const Vol v{16, w_.ir.height, 16, cx_ * 16, ymin_, cz_ * 16, 1, 1, 1};
            std::vector<float> b(static_cast<size_t>(v.n()));
            d_.sample(w_.ir.roots.at(root), v, b.data());
            it = ore_.emplace(root, std::move(b)).first;
        }
        const Vol v{16, w_.ir.height, 16, cx_ * 16, ymin_, cz_ * 16, 1, 1, 1};
        return it->second[static_cast<size_t>(v.at(lx_, by_ - ymin_, lz_))];
    }

    int rule(int ri) {
        const Rule& r = m_.rules[static_cast<size_t>(ri)];
        if (r.t == "block") return r.state;
        if (r.t == "sequence") {
            for (int s : r.seq) {
                const int v = rule(s);
                if (v >= 0) return v;
            }
            return -1;
        }
        if (r.t == "condition") return cond(r.cond) ? rule(r.then) : -1;
        if (r.t == "bandlands") {
            const float off = static_cast<float>(nz("minecraft:clay_bands_offset", bx_, 0.0, bz_)) * 4.0f;
            const int o = to_int(std::floor(static_cast<double>(off) + 0.5));   // Math.round(float)
            return w_.clay[(by_ + o + 192) % 192];
        }
        if (r.t == "ore_vein") {
            const float density = ore_value(r.density);
            if (density <= 0.0f) return -1;
            Xrand rnd = w_.ore.at(bx_, by_, bz_);
            if (rnd.next_float() > density) return -1;
            const float rich = ore_value(r.richness);
            if (rnd.next_float() < rich && d_.point(w_.ir.roots.at(r.gap), bx_, by_, bz_) < 0.0f) return rnd.next_float() < r.raw_chance ? r.raw : r.ore;
            return r.filler;
        }
        throw std::runtime_error("rule " + r.t);
    }

    void eroded_badlands(int bx, int bz, int height) {
        const double a = std::fabs(nz("minecraft:badlands_surface", bx, 0.0, bz) * 8.25);
        const double b = static_cast<double>(static_cast<float>(nz("minecraft:badlands_pillar", bx * 0.2, 0.0, bz * 0.2)) * 15.0f);
        const double pillar = java_min(a, b);
        if (pillar <= 0.0) return;
        const double roof = std::fabs(nz("minecraft:badlands_pillar_roof", bx * 0.75, 0.0, bz * 0.75) * 1.5);
        const int top = floor_i(64.0 + java_min(pillar * pillar * 2.5, std::ceil(roof * 50.0) + 24.0));
        if (height > top) return;
        for (int y = top; y >= ymin_; --y) {
            const int s = get(y);
            if (s == m_.default_block) break;
            if (m_.is_water[static_cast<size_t>(s)]) return;
        }
        for (int y = top; y >= ymin_ && air(get(y)); --y) set(y, m_.default_block);
    }

    void frozen_ocean(int min_surface, int biome, int bx, int bz, int height) {
        const double a = std::fabs(nz("minecraft:iceberg_surface", bx, 0.0, bz) * 8.25);
        const double b = static_cast<double>(static_cast<float>(nz("minecraft:iceberg_pillar", bx * 1.28, 0.0, bz * 1.28)) * 15.0f);
        const double berg = java_min(a, b);
        if (berg <= 1.8) return;
        const double roof = std::fabs(nz("minecraft:iceberg_pillar_roof", bx * 1.17, 0.0, bz * 1.17) * 1.5);
        double top = java_min(berg * berg * 1.2, std::ceil(roof * 40.0) + 14.0);
        const int sea = m_.sea;
        if (temperature(biome, bx, sea, bz) > 0.1f) top -= 2.0;
        if (top <= 2.0) return;
        const double bottom = static_cast<double>(sea) - top - 7.0;
        top += static_cast<double>(sea);
        Xrand r = w_.random.at(bx, 0, bz);
        const int max_snow = 2 + r.next_int(4);
        const int snow_from = sea + 18 + r.next_int(10);
        const int itop = to_int(top), ibottom = to_int(bottom);
        int snow = 0;
        for (int y = std::max(height, itop + 1); y >= min_surface; --y) {
            const bool place = (air(get(y)) && y < itop && r.next_double() > 0.01) ||
                               (get(y) == m_.water && y > ibottom && y < sea && r.next_double() > 0.15);
            if (!place) continue;
            if (snow <= max_snow && y > snow_from) {
                set(y, m_.snow);
                ++snow;
            } else {
                set(y, m_.packed_ice);
            }
        }
    }

    World& w_;
    Density& d_;
    const Model& m_;
    int cx_, cz_;
    uint16_t* blocks_;
    int* hm_;
    const QuartBiome& qb_;
    int ymin_, ymax_;
    std::vector<float> prelim_;
    std::map<std::string, std::vector<float>> ore_;
    int lx_ = 0, lz_ = 0, bx_ = 0, by_ = 0, bz_ = 0, gx_ = 0, gz_ = 0, depth_ = 0, above_ = 0, below_ = 0, water_ = 0;
    bool have_secondary_ = false, have_min_surface_ = false, have_biome_ = false;
    double secondary_ = 0;
    int min_surface_ = 0, biome_ = 0;
};

}  // namespace

World::World(const Ir& ir_, const Model& m_, int64_t seed_)
    : ir(ir_), m(m_), seed(seed_), tables(make_tables(ir_, seed_)), tree(m_.params), zseed(zoom_seed_of(seed_)) {
    Xrand base = Xrand::from_seed(seed_);
    random = Factory::fork(base);
    Xrand ra = random.hash("minecraft:aquifer");
    aquifer = Factory::fork(ra);
    Xrand ro = random.hash("minecraft:ore");
    ore = Factory::fork(ro);
    for (const Cond& c : m.conds)
        if (c.t == "vertical_gradient" && !gradient.count(c.random_name)) {
            Xrand rg = random.hash(c.random_name);
            gradient.emplace(c.random_name, Factory::fork(rg));
        }
    // MaterialSystem.generateBands (terrain.cpp:117-149)
    Xrand r = random.hash("minecraft:clay_bands");
    const int* tc = m.terracotta;   // plain, white, orange, yellow, brown, red, light gray
    std::fill(clay, clay + 192, tc[0]);
    for (int i = 0; i < 192; ++i) {
        i += r.next_int(5) + 1;
        if (i < 192) clay[i] = tc[2];
    }
    const int widths[3] = {1, 2, 1}, states[3] = {tc[3], tc[4], tc[5]};
    for (int k = 0; k < 3; ++k) {
        const int count = r.next_int(10) + 6;
        for (int i = 0; i < count; ++i) {
            const int width = widths[k] + r.next_int(3), start = r.next_int(192);
            for (int p = 0; start + p < 192 && p < width; ++p) clay[start + p] = states[k];
        }
    }
    const int whites = r.next_int(7) + 9;
    for (int i = 0, start = 0; i < whites && start < 192; start += r.next_int(16) + 4, ++i) {
        clay[start] = tc[1];
        if (start - 1 > 0 && r.next_bool()) clay[start - 1] = tc[6];
        if (start + 1 < 192 && r.next_bool()) clay[start + 1] = tc[6];
    }
    // Biome temperature noises: LegacyRandomSource 1234 (temperature), 3456 (frozen, three
    // layers), 2345 (biome info)
    Lcg t(1234), f(3456), b(2345);
    simplex_perm(t, simplex[0]);
    for (int k = 0; k < 3; ++k) simplex_perm(f, simplex[1 + k]);
    simplex_perm(b, simplex[4]);
}

// Biome.getHeightAdjustedTemperature (terrain.cpp:87-110)
float biome_temperature(const World& w, int biome, int x, int y, int z) {
    float t = w.m.biome_temp[static_cast<size_t>(biome)];
    if (w.m.biome_frozen[static_cast<size_t>(biome)]) {
        const float amp[3] = {0.14285715f, 0.2857143f, 0.5714286f};
        const double freq[3] = {1.0, 0.5, 0.25};
        float frozen = 0.0f;
        for (int k = 0; k < 3; ++k) frozen += amp[k] * simplex(w.simplex[1 + k], x * 0.05 * freq[k], z * 0.05 * freq[k]);
        const double patches = static_cast<double>(frozen * 7.0f) + static_cast<double>(simplex(w.simplex[4], x * 0.2, z * 0.2));
        if (patches < 0.3 && static_cast<double>(simplex(w.simplex[4], x * 0.09, z * 0.09)) < 0.8) t = 0.2f;
    }
    const int snow = w.m.sea + 17;
    if (y > snow) {
        const float n = simplex(w.simplex[0], static_cast<double>(static_cast<float>(x) / 8.0f), static_cast<double>(static_cast<float>(z) / 8.0f)) * 8.0f;
        return t - ((n + static_cast<float>(y) - static_cast<float>(snow)) * 0.05f) / 40.0f;
    }
    return t;
}

// Quart biomes of one chunk: the six climate channels over a 4x(h/4)x4 quart volume, then the
// R-tree search with the previous result as candidate, in vanilla's fill order (terrain.cpp:152-178)
void chunk_biomes(World& w, Density& d, int cx, int cz, uint8_t* out) {
    static const char* roots[6] = {"climate/temperature", "climate/vegetation", "climate/continents", "climate/erosion", "climate/depth", "climate/ridges"};
    const int qh = w.ir.height / 4;
    const Vol v{4, qh, 4, cx * 16, w.ir.min_y, cz * 16, 4, 4, 4};
    std::vector<float> ch[6];
    d.reset();
    for (int k = 0; k < 6; ++k) {
        ch[k].resize(static_cast<size_t>(v.n()));
        d.sample(w.ir.roots.at(roots[k]), v, ch[k].data());
    }
    int last = -1;
    for (int sy = 0; sy < qh / 4; ++sy)
        for (int x = 0; x < 4; ++x)
            for (int y = 0; y < 4; ++y)
                for (int z = 0; z < 4; ++z) {
                    const int qy = sy * 4 + y, i = v.at(x, qy, z);
                    int64_t t[7];
                    for (int k = 0; k < 6; ++k) t[k] = quantize(ch[k][static_cast<size_t>(i)]);
                    t[6] = 0;
                    last = w.tree.search(t, last);
                    out[(qy * 4 + z) * 4 + x] = static_cast<uint8_t>(w.m.params[static_cast<size_t>(last)].biome);
                }
}

// Final density, aquifer fill (z, x, y descending) and surface of one chunk (terrain.cpp:779-810)
void chunk_terrain(World& w, Density& d, int cx, int cz, const uint8_t* const nb[9], uint16_t* aquifer, uint16_t* terrain, float* density) {
    const int h = w.ir.height, qh = h / 4, qmin = w.ir.min_y >> 2;
    const Model& m = w.m;
    const Surface::QuartBiome qb = [&](int qx, int qy, int qz) {
        const int dx = (qx >> 2) - cx, dz = (qz >> 2) - cz;
        if (dx < -1 || dx > 1 || dz < -1 || dz > 1) throw std::runtime_error("biome lookup outside the 3x3 neighbourhood");
        const int y = std::min(std::max(qy, qmin), qmin + qh - 1) - qmin;
        return static_cast<int>(nb[(dz + 1) * 3 + dx + 1][(y * 4 + (qz & 3)) * 4 + (qx & 3)]);
    };
    const Vol v{16, h, 16, cx * 16, w.ir.min_y, cz * 16, 1, 1, 1};
    d.reset();
    Aquifer aq(w, d, v);
    std::vector<float> dens(static_cast<size_t>(v.n()));
    d.sample(w.ir.roots.at("final_density"), v, dens.data());
    std::vector<uint16_t> fill(static_cast<size_t>(v.n()));
    int hm[256];
    std::fill(hm, hm + 256, w.ir.min_y);
    for (int z = 0; z < 16; ++z)
        for (int x = 0; x < 16; ++x)
            for (int y = h - 1; y >= 0; --y) {
                const float dv = dens[static_cast<size_t>(v.at(x, y, z))];
                const size_t o = static_cast<size_t>((y * 16 + z) * 16 + x);
                if (density) density[o] = dv;
                int s = aq.substance(cx * 16 + x, w.ir.min_y + y, cz * 16 + z, static_cast<double>(dv));
                if (s < 0) s = m.default_block;
                fill[o] = static_cast<uint16_t>(s);
                if (!m.is_air[static_cast<size_t>(s)]) {
                    int& top = hm[x * 16 + z];
                    const int by = w.ir.min_y + y;
                    if (!(by <= top - 2) && by >= top) top = by + 1;
                }
            }
    if (aquifer) std::copy(fill.begin(), fill.end(), aquifer);
    Surface surf(w, d, cx, cz, fill.data(), hm, qb);
    surf.run();
    if (terrain) std::copy(fill.begin(), fill.end(), terrain);
}

}  // namespace orc
