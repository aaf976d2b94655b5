// Oracle: readers of the IR (mccuda-ir/1) and terrain model (mccuda-terrain/1) files, the
// per-seed noise tables, and the climate R-tree with the BiomeManager zoom.
// Restates cpu/src/ir.cpp:63-157, cpu/src/terrain_model.cpp:111-169,
// cpu/src/seed_tables.cpp:23-184 and cpu/src/biome.cpp:61-289.
#include <algorithm>
#include <limits>
#include <stdexcept>

#include "oracle.h"

namespace orc {

// ---- IR -----------------------------------------------------------------------------------------
namespace {
SplineN read_spline(const Json& j) {
    SplineN s;
    if (j.has("k")) {
        s.is_const = true;
        s.k = j["k"].f32bits();
        return s;
    }
    s.coord = static_cast<int>(j["coord"].i64());
    for (const Json& v : j["loc"].a) s.loc.push_back(v.f32bits());
    for (const Json& v : j["der"].a) s.der.push_back(v.f32bits());
    for (const Json& v : j["val"].a) s.val.push_back(read_spline(v));
    return s;
}
}  // namespace

Ir load_ir(const std::string& path) {
    const Json doc = parse_json_file(path);
    if (doc["format"].s != "mccuda-ir/1") throw std::runtime_error("not an mccuda-ir/1 file");
    Ir ir;
    ir.min_y = static_cast<int>(doc["min_y"].i64());
    ir.height = static_cast<int>(doc["height"].i64());
    ir.sea = static_cast<int>(doc["sea_level"].i64());
    for (const auto& kv : doc["roots"].o) ir.roots[kv.first] = static_cast<int>(kv.second.i64());
    const std::map<std::string, T> types = {
        {"cache", T::Cache}, {"constant", T::Const}, {"simple", T::Simple}, {"gradient", T::Gradient}, {"noise", T::Noise},
        {"shift_a", T::ShiftA}, {"shift_b", T::ShiftB}, {"shift", T::Shift}, {"blended_noise", T::Blended}, {"binary", T::Binary},
        {"unary", T::Unary}, {"clamp", T::Clamp}, {"lerp", T::Lerp}, {"range_choice", T::RangeChoice}, {"interval_select", T::IntervalSelect},
        {"spline", T::Spline}, {"interpolated", T::Interp}, {"slice", T::Slice}, {"blend_density", T::BlendDensity}, {"find_top_surface", T::Fts}};
    for (const Json& j : doc["nodes"].a) {
        Node n;
        if (static_cast<size_t>(j["id"].i64()) != ir.nodes.size()) throw std::runtime_error("node ids must be dense and in order");
        const auto it = types.find(j["t"].s);
        if (it == types.end()) throw std::runtime_error("unknown node type " + j["t"].s);
        n.t = it->second;
        n.lo = j["range"][0].f32bits();
        n.hi = j["range"][1].f32bits();
        auto gi = [&](const char* k) { return static_cast<int>(j[k].i64()); };
        switch (n.t) {
            case T::Cache: case T::BlendDensity: n.in = gi("in"); break;
            case T::Const: n.v = j["v"].f32bits(); break;
            case T::Simple: n.kind = j["kind"].s; break;
            case T::Gradient:
                n.axis = j["axis"].s; n.tiling = j["tiling"].s; n.from = gi("from"); n.to = gi("to");
                n.fv = j["from_value"].f32bits(); n.tv = j["to_value"].f32bits();
                break;
            case T::Noise:
                n.noise = j["noise"].s; n.xz = j["xz_scale"].hexd(); n.ys = j["y_scale"].hexd();
                n.sx = gi("shift_x"); n.sy = gi("shift_y"); n.sz = gi("shift_z");
                n.zx = j["shift_x_zero"].b; n.zy = j["shift_y_zero"].b; n.zz = j["shift_z_zero"].b;
                break;
            case T::ShiftA: case T::ShiftB: case T::Shift: n.noise = j["noise"].s; break;
            case T::Blended:
                n.xz = j["xz_scale"].hexd(); n.ys = j["y_scale"].hexd(); n.xzf = j["xz_factor"].hexd(); n.yf = j["y_factor"].hexd();
                n.smear = j["smear_scale_multiplier"].hexd();
                break;
            case T::Binary: n.op = j["op"].s; n.l = gi("l"); n.r = gi("r"); break;
            case T::Unary: n.op = j["op"].s; n.in = gi("in"); break;
            case T::Clamp: n.in = gi("in"); n.fmin = j["min"].f32bits(); n.fmax = j["max"].f32bits(); break;
            case T::Lerp: n.alpha = gi("alpha"); n.first = gi("first"); n.second = gi("second"); break;
            case T::RangeChoice:
                n.in = gi("in"); n.fmin = j["min_inclusive"].f32bits(); n.fmax = j["max_exclusive"].f32bits();
                n.when_in = gi("when_in"); n.when_out = gi("when_out");
                break;
            case T::IntervalSelect:
                n.in = gi("in");
                for (const Json& v : j["thresholds"].a) n.th.push_back(v.f32bits());
                for (const Json& v : j["functions"].a) n.fns.push_back(static_cast<int>(v.i64()));
                if (n.th.size() + 1 != n.fns.size()) throw std::runtime_error("interval_select arity");
                break;
            case T::Spline: n.spline = std::make_shared<SplineN>(read_spline(j["spline"])); break;
            case T::Interp: n.in = gi("in"); n.cell_xz = gi("cell_xz"); n.cell_y = gi("cell_y"); break;
            case T::Slice: n.axis = j["axis"].s; n.coord = gi("coord"); n.in = gi("in"); break;
            case T::Fts: n.density = gi("density"); n.upper = gi("upper_bound"); n.lower = gi("lower_bound"); n.cell_h = gi("cell_height"); break;
        }
        ir.nodes.push_back(std::move(n));
    }
    for (const auto& kv : doc["noises"].o) {
        NoiseParams p;
        p.base_amp = kv.second["base_amplitude"].hexd();
        p.base_octave = static_cast<int>(kv.second["base_octave"].i64());
        p.count = static_cast<int>(kv.second["octave_count"].i64());
        p.normalize = kv.second["normalize"].s;
        for (const Json& m : kv.second["amplitude_modifiers"].a) p.mods.push_back(m.hexd());
        ir.noises[kv.first] = p;
    }
    return ir;
}

// ---- terrain model ------------------------------------------------------------------------------
namespace {
struct ModelReader {
    Model& m;
    int cond(const Json& j) {
        if (j["t"].s == "ref") return cond(j["cond"]);
        Cond c;
        c.t = j["t"].s;
        if (c.t == "biome") {
            for (const Json& b : j["biomes"].a) {
                const int64_t i = b.i64();
                if (i >= 128) throw std::runtime_error("biome index >= 128");
                c.mask[i >> 6] |= 1ull << (i & 63);
            }
        } else if (c.t == "noise_threshold") {
            c.noise = j["noise"].s; c.mn = j["min"].hexd(); c.mx = j["max"].hexd(); c.is3d = j["is_3d"].b;
        } else if (c.t == "vertical_gradient") {
            c.random_name = j["random_name"].s;
            c.true_at = static_cast<int>(j["true_at_and_below"].i64());
            c.false_at = static_cast<int>(j["false_at_and_above"].i64());
        } else if (c.t == "y_above") {
            c.anchor = static_cast<int>(j["anchor"].i64()); c.mult = static_cast<int>(j["surface_depth_multiplier"].i64()); c.add_stone = j["add_stone_depth"].b;
        } else if (c.t == "water") {
            c.offset = static_cast<int>(j["offset"].i64()); c.mult = static_cast<int>(j["surface_depth_multiplier"].i64()); c.add_stone = j["add_stone_depth"].b;
        } else if (c.t == "not") {
            c.child = cond(j["cond"]);
        } else if (c.t == "stone_depth") {
            c.offset = static_cast<int>(j["offset"].i64()); c.add_surface = j["add_surface_depth"].b;
            c.secondary = static_cast<int>(j["secondary_depth_range"].i64()); c.ceiling = j["ceiling"].b;
        } else if (c.t != "temperature" && c.t != "steep" && c.t != "hole" && c.t != "above_preliminary_surface") {
            throw std::runtime_error("unknown condition " + c.t);
        }
        m.conds.push_back(c);
        return static_cast<int>(m.conds.size()) - 1;
    }
    int rule(const Json& j) {
        if (j["t"].s == "ref") return rule(j["rule"]);
        Rule r;
        r.t = j["t"].s;
        if (r.t == "block") r.state = static_cast<int>(j["state"].i64());
        else if (r.t == "sequence") for (const Json& x : j["rules"].a) r.seq.push_back(rule(x));
        else if (r.t == "condition") { r.cond = cond(j["if"]); r.then = rule(j["then"]); }
        else if (r.t == "ore_vein") {
            r.ore = static_cast<int>(j["ore"].i64()); r.raw = static_cast<int>(j["raw_ore"].i64()); r.filler = static_cast<int>(j["filler"].i64());
            r.raw_chance = j["raw_ore_chance"].f32bits();
            r.density = j["density"].s; r.richness = j["richness"].s; r.gap = j["filler_gap"].s;
        } else if (r.t != "bandlands") throw std::runtime_error("unknown rule " + r.t);
        m.rules.push_back(r);
        return static_cast<int>(m.rules.size()) - 1;
    }
};
}  // namespace

Model load_model(const std::string& path) {
    const Json doc = parse_json_file(path);
    if (doc["format"].s != "mccuda-terrain/1") throw std::runtime_error("not an mccuda-terrain/1 file");
    Model m;
    m.sea = static_cast<int>(doc["sea_level"].i64());
    for (const Json& p : doc["palette"].a) {
        m.palette.push_back(p.s);
        m.is_air.push_back(p.s == "minecraft:air");
        m.is_water.push_back(p.s.rfind("minecraft:water", 0) == 0);
        m.is_lava.push_back(p.s.rfind("minecraft:lava", 0) == 0);
    }
    auto find = [&](const std::string& n) {
        for (size_t i = 0; i < m.palette.size(); ++i)
            if (m.palette[i] == n) return static_cast<int>(i);
        return -1;
    };
    m.default_block = static_cast<int>(doc["default_block"].i64());
    m.default_fluid = static_cast<int>(doc["default_fluid"].i64());
    m.air = find("minecraft:air");
    m.water = find("minecraft:water[level=0]");
    m.lava = find("minecraft:lava[level=0]");
    m.packed_ice = find("minecraft:packed_ice");
    m.snow = find("minecraft:snow_block");
    const char* tc[7] = {"terracotta", "white_terracotta", "orange_terracotta", "yellow_terracotta", "brown_terracotta", "red_terracotta",
                         "light_gray_terracotta"};
    for (int i = 0; i < 7; ++i) m.terracotta[i] = find(std::string("minecraft:") + tc[i]);
    if (m.air < 0 || m.water < 0 || m.lava < 0 || m.packed_ice < 0 || m.snow < 0 || m.terracotta[0] < 0 || m.terracotta[6] < 0)
        throw std::runtime_error("palette lacks a required block state");
    for (const Json& n : doc["biome_names"].a) m.biome_names.push_back(n.s);
    const Json& bl = doc["biomes"];
    for (size_t i = 0; i < m.biome_names.size(); ++i) {
        if (bl[i]["name"].s != m.biome_names[i]) throw std::runtime_error("biome list order");
        m.biome_temp.push_back(bl[i]["temperature"].f32bits());
        m.biome_frozen.push_back(bl[i]["temperature_modifier"].s == "frozen");
        if (m.biome_names[i] == "minecraft:eroded_badlands") m.eroded_badlands = static_cast<int>(i);
        if (m.biome_names[i] == "minecraft:frozen_ocean") m.frozen_ocean = static_cast<int>(i);
        if (m.biome_names[i] == "minecraft:deep_frozen_ocean") m.deep_frozen_ocean = static_cast<int>(i);
    }
    for (const Json& p : doc["biome_parameters"].a) {
        Box b;
        for (int d = 0; d < 6; ++d) {
            b.mn[d] = p["p"][2 * d].i64();
            b.mx[d] = p["p"][2 * d + 1].i64();
        }
        b.mn[6] = b.mx[6] = p["offset"].i64();
        b.biome = static_cast<int>(p["biome"].i64());
        m.params.push_back(b);
    }
    ModelReader r{m};
    m.root = r.rule(doc["material_rule"]);
    return m;
}

// ---- per-seed noise tables ------------------------------------------------------------------------
double stream_sum(const std::vector<double>& v) {   // DoubleStream.sum: Kahan summation
    double sum = 0.0, comp = 0.0, plain = 0.0;
    for (double x : v) {
        const double y = x - comp;
        const double t = sum + y;
        comp = (t - sum) - y;
        sum = t;
        plain += x;
    }
    const double r = sum - comp;
    return (std::isnan(r) && std::isinf(plain)) ? plain : r;
}

namespace {
Octave draw_octave(Xrand& r, double smear) {   // GradientNoise(random, 256.0)
    Octave o{};
    o.xo = r.next_double() * 256.0;
    o.yo = r.next_double() * 256.0;
    o.zo = r.next_double() * 256.0;
    o.smear = smear;
    for (int i = 0; i < 256; ++i) o.perm[i] = static_cast<uint8_t>(i);
    for (int i = 0; i < 256; ++i) std::swap(o.perm[i], o.perm[i + r.next_int(256 - i)]);
    return o;
}

// NormalNoise(params).create(random): two octave sets from two positional forks
Stack normal_noise(Tables& t, const NoiseParams& p, Xrand random) {
    struct Oct { int index; double freq, amp; };
    std::vector<Oct> octs;
    auto mod = [&](int i) { return p.mods.empty() ? 1.0 : p.mods[static_cast<size_t>(i)]; };
    double freq = std::pow(2.0, p.base_octave), amp = p.base_amp;
    if (p.normalize != "disabled") amp *= std::pow(2.0, p.count - 1) / (std::pow(2.0, p.count) - 1.0);
    for (int i = 0; i < p.count; ++i) {
        if (mod(i) != 0.0) octs.push_back({p.base_octave + i, freq, amp * mod(i)});
        freq *= 2.0;
        amp *= 0.5;
    }
    double variance = 0.0;
    std::vector<double> abs_amp;
    for (const Oct& o : octs) {
        const double dev = 0.2702247831245211 * std::fabs(o.amp);
        variance += dev * dev;
        abs_amp.push_back(std::fabs(o.amp));
    }
    const double deviation = std::sqrt(variance);
    double norm = 0.0;
    if (deviation != 0.0) norm = (stream_sum(abs_amp) * 0.3333333333333333) / (deviation * std::sqrt(2.0));
    if (p.normalize == "legacy" && norm != 0.0) {
        int lo = INT32_MAX, hi = INT32_MIN;
        for (int i = 0; i < p.count; ++i)
            if (mod(i) != 0.0) { lo = std::min(lo, i); hi = std::max(hi, i); }
        norm = p.base_amp * 0.5 * 0.3333333333333333 / (0.1 * (1.0 + 1.0 / (hi - lo + 1)));
    }
    const Factory a = Factory::fork(random);
    const Factory b = Factory::fork(random);
    Stack s;
    s.first = static_cast<int>(t.layers.size());
    for (const Oct& o : octs) {
        const std::string name = "octave_" + std::to_string(o.index);
        Xrand ra = a.hash(name), rb = b.hash(name);
        const float vf = static_cast<float>(norm * o.amp);
        t.oct.push_back(draw_octave(ra, 0.0));
        t.layers.push_back({o.freq, vf, static_cast<int>(t.oct.size()) - 1});
        t.oct.push_back(draw_octave(rb, 0.0));
        t.layers.push_back({o.freq * 1.0181268882175227, vf, static_cast<int>(t.oct.size()) - 1});
    }
    s.count = static_cast<int>(t.layers.size()) - s.first;
    return s;
}

// BlendedNoise.createFbm(random, firstOctave): coarsest frequency 1, halving per octave
Stack fbm(Tables& t, Xrand& r, int first_octave, double smear_y, double value_factor) {
    const int n = 1 - first_octave;
    value_factor /= std::pow(2.0, n) - 1.0;
    double factor = 1.0;
    Stack s;
    s.first = static_cast<int>(t.layers.size());
    s.smeared = true;
    for (int i = 0; i < n; ++i) {
        t.oct.push_back(draw_octave(r, smear_y * factor));
        t.layers.push_back({factor, static_cast<float>(value_factor), static_cast<int>(t.oct.size()) - 1});
        factor /= 2.0;
        value_factor *= 2.0;
    }
    s.count = static_cast<int>(t.layers.size()) - s.first;
    return s;
}
}  // namespace

Tables make_tables(const Ir& ir, int64_t seed) {
    Tables t;
    Xrand base = Xrand::from_seed(seed);
    const Factory world = Factory::fork(base);
    for (const auto& kv : ir.noises) t.noise[kv.first] = normal_noise(t, kv.second, world.hash(kv.first));
    for (size_t id = 0; id < ir.nodes.size(); ++id) {
        const Node& n = ir.nodes[id];
        if (n.t != T::Blended) continue;
        Xrand r = world.hash("minecraft:terrain");
        const double limit_smear = 684.412 * n.ys * n.smear;
        Stack (&b)[3] = t.blended[static_cast<int>(id)];
        b[0] = fbm(t, r, -15, limit_smear, static_cast<double>(0.99998474f));
        b[1] = fbm(t, r, -15, limit_smear, static_cast<double>(0.99998474f));
        b[2] = fbm(t, r, -7, limit_smear / n.yf, 12.75);
    }
    return t;
}

float stack_get(const Tables& t, const Stack& s, double x, double y, double z) {
    float v = 0.0f;
    for (int i = 0; i < s.count; ++i) {
        const Layer& L = t.layers[static_cast<size_t>(s.first + i)];
        const Octave& o = t.oct[static_cast<size_t>(L.octave)];
        const float n = s.smeared ? noise_get_smeared(o, x * L.freq, y * L.freq, z * L.freq) : noise_get(o, x * L.freq, y * L.freq, z * L.freq);
        v += L.amp * n;
    }
    return v;
}

// ---- climate R-tree (Climate.RTree) -----------------------------------------------------------------
namespace {
int64_t mid(int64_t a, int64_t b) { return (a + b) / 2; }   // Java long division truncates
int64_t dim_dist(int64_t mn, int64_t mx, int64_t t) {
    const int64_t above = static_cast<int64_t>(static_cast<uint64_t>(t) - static_cast<uint64_t>(mx));
    const int64_t below = static_cast<int64_t>(static_cast<uint64_t>(mn) - static_cast<uint64_t>(t));
    return above > 0 ? above : std::max<int64_t>(below, 0);
}
}  // namespace

int64_t quantize(float v) {
    const float q = v * 10000.0f;
    if (std::isnan(q)) return 0;
    if (q >= 9.2233720368547758e18f) return INT64_MAX;
    if (q <= -9.2233720368547758e18f) return INT64_MIN;
    return static_cast<int64_t>(q);
}

RTree::RTree(const std::vector<Box>& params) {
    std::vector<int> items;
    for (size_t i = 0; i < params.size(); ++i) {
        N n;
        std::memcpy(n.mn, params[i].mn, sizeof n.mn);
        std::memcpy(n.mx, params[i].mx, sizeof n.mx);
        n.leaf = static_cast<int>(i);
        n_.push_back(n);
        leaf_node_.push_back(static_cast<int>(i));
        items.push_back(static_cast<int>(i));
    }
    root_ = build(items);
}

int RTree::add(const std::vector<int>& kids) {
    N n;
    for (int d = 0; d < 7; ++d) {
        n.mn[d] = INT64_MAX;
        n.mx[d] = INT64_MIN;
        for (int k : kids) {
            n.mn[d] = std::min(n.mn[d], n_[static_cast<size_t>(k)].mn[d]);
            n.mx[d] = std::max(n.mx[d], n_[static_cast<size_t>(k)].mx[d]);
        }
    }
    n.kids = kids;
    n_.push_back(n);
    return static_cast<int>(n_.size()) - 1;
}

int RTree::build(std::vector<int> items) {
    if (items.size() == 1) return items[0];
    auto c = [&](int k, int d) { return mid(n_[static_cast<size_t>(k)].mn[d], n_[static_cast<size_t>(k)].mx[d]); };
    if (items.size() <= 19) {   // leaf-level node: order by the sum of |centres|
        std::stable_sort(items.begin(), items.end(), [&](int a, int b) {
            int64_t sa = 0, sb = 0;
            for (int d = 0; d < 7; ++d) { sa += std::llabs(c(a, d)); sb += std::llabs(c(b, d)); }
            return sa < sb;
        });
        return add(items);
    }
    const int size = static_cast<int>(std::pow(19.0, std::floor(std::log(static_cast<double>(items.size()) - 0.01) / std::log(19.0))));
    auto order = [&](std::vector<int>& v, int dim, bool absolute) {
        std::stable_sort(v.begin(), v.end(), [&](int a, int b) {
            for (int k = 0; k < 7; ++k) {
                const int d = (dim + k) % 7;
                int64_t ca = c(a, d), cb = c(b, d);
                if (absolute) { ca = std::llabs(ca); cb = std::llabs(cb); }
                if (ca != cb) return ca < cb;
            }
            return false;
        });
    };
    int64_t best = INT64_MAX;
    int best_dim = 0;
    std::vector<std::vector<int>> best_groups;
    for (int d = 0; d < 7; ++d) {
        order(items, d, false);
        std::vector<std::vector<int>> groups;
        for (size_t i = 0; i < items.size(); i += static_cast<size_t>(size))
            groups.emplace_back(items.begin() + static_cast<long>(i), items.begin() + static_cast<long>(std::min(items.size(), i + size)));
        int64_t cost = 0;
        for (const auto& g : groups)
            for (int dd = 0; dd < 7; ++dd) {
                int64_t mn = INT64_MAX, mx = INT64_MIN;
                for (int k : g) { mn = std::min(mn, n_[static_cast<size_t>(k)].mn[dd]); mx = std::max(mx, n_[static_cast<size_t>(k)].mx[dd]); }
                cost += std::llabs(mx - mn);
            }
        if (cost < best) { best = cost; best_dim = d; best_groups = groups; }
    }
    std::vector<int> holders;
    for (const auto& g : best_groups) holders.push_back(add(g));
    order(holders, best_dim, true);
    std::vector<int> kids;
    for (int h : holders) kids.push_back(build(n_[static_cast<size_t>(h)].kids));
    return add(kids);
}

int64_t RTree::dist(int n, const int64_t t[7]) const {
    uint64_t s = 0;
    for (int d = 0; d < 7; ++d) {
        const int64_t p = dim_dist(n_[static_cast<size_t>(n)].mn[d], n_[static_cast<size_t>(n)].mx[d], t[d]);
        s += static_cast<uint64_t>(p) * static_cast<uint64_t>(p);
    }
    return static_cast<int64_t>(s);
}

int64_t RTree::leaf_dist(int leaf, const int64_t t[7]) const { return dist(leaf_node_[static_cast<size_t>(leaf)], t); }

// SubTree.search(target, candidate): descend only into strictly closer boxes, keep a leaf only
// if strictly closer than the best so far (which starts at the candidate).
int RTree::walk(int n, const int64_t t[7], int cand) const {
    const N& node = n_[static_cast<size_t>(n)];
    if (node.leaf >= 0) return n;
    int best = cand;
    int64_t best_d = cand < 0 ? INT64_MAX : dist(cand, t);
    for (int k : node.kids) {
        const int64_t dk = dist(k, t);
        if (best_d <= dk) continue;
        const int found = walk(k, t, best);
        const int64_t df = found == k ? dk : dist(found, t);
        if (df < best_d) { best_d = df; best = found; }
    }
    return best;
}

int RTree::search(const int64_t t[7], int candidate_leaf) const {
    const int cand = candidate_leaf < 0 ? -1 : leaf_node_[static_cast<size_t>(candidate_leaf)];
    return n_[static_cast<size_t>(walk(root_, t, cand))].leaf;
}

int RTree::brute_first_min(const int64_t t[7], int64_t* d) const {
    // every leaf in DFS order; the first one at the minimal distance
    std::vector<int> order;
    std::vector<int> st{root_};
    while (!st.empty()) {
        const int n = st.back();
        st.pop_back();
        if (n_[static_cast<size_t>(n)].leaf >= 0) { order.push_back(n); continue; }
        const auto& k = n_[static_cast<size_t>(n)].kids;
        for (auto it = k.rbegin(); it != k.rend(); ++it) st.push_back(*it);
    }
    int best = -1;
    int64_t bd = INT64_MAX;
    for (int n : order) {
        const int64_t dn = dist(n, t);
        if (best < 0 || dn < bd) { best = n; bd = dn; }
    }
    if (d) *d = bd;
    return n_[static_cast<size_t>(best)].leaf;
}

// ---- BiomeManager zoom -------------------------------------------------------------------------------
namespace {
int64_t lcg_step(int64_t r, int64_t c) {
    const uint64_t u = static_cast<uint64_t>(r);
    return static_cast<int64_t>(u * (u * 6364136223846793005ull + 1442695040888963407ull) + static_cast<uint64_t>(c));
}
double fiddle(int64_t r) {
    int64_t m = (r >> 24) % 1024;
    if (m < 0) m += 1024;
    return (static_cast<double>(m) / 1024.0 - 0.5) * 0.9;
}
}  // namespace

void zoom_quart(int64_t zs, int x, int y, int z, int& qx, int& qy, int& qz) {
    const int ax = x - 2, ay = y - 2, az = z - 2;
    const int px = ax >> 2, py = ay >> 2, pz = az >> 2;
    const double fx = (ax & 3) / 4.0, fy = (ay & 3) / 4.0, fz = (az & 3) / 4.0;
    int best = 0;
    double best_d = std::numeric_limits<double>::infinity();
    for (int i = 0; i < 8; ++i) {
        const int cx = px + ((i >> 2) & 1), cy = py + ((i >> 1) & 1), cz = pz + (i & 1);
        const double dx = fx - ((i >> 2) & 1), dy = fy - ((i >> 1) & 1), dz = fz - (i & 1);
        int64_t r = zs;
        for (int rep = 0; rep < 2; ++rep) { r = lcg_step(r, cx); r = lcg_step(r, cy); r = lcg_step(r, cz); }
        const double ox = fiddle(r);
        r = lcg_step(r, zs);
        const double oy = fiddle(r);
        r = lcg_step(r, zs);
        const double oz = fiddle(r);
        const double d = (dz + oz) * (dz + oz) + (dy + oy) * (dy + oy) + (dx + ox) * (dx + ox);
        if (d < best_d) { best_d = d; best = i; }
    }
    qx = px + ((best >> 2) & 1);
    qy = py + ((best >> 1) & 1);
    qz = pz + (best & 1);
}

}  // namespace orc
