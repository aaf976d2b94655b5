// Oracle: the density-function interpreter. Volume path = DensitySampler.sampleVolume over a
// block volume (buffers, cell interpolation, slices, caches keyed by volume); point path =
// DensitySampler.sampleValue. Restates cpu/src/volume_eval.cpp:38-674, including vanilla's
// compile-time binary-operator decisions (constant operands, division by a constant as a
// multiply by its float reciprocal, MIN/MAX pruning from value ranges).
#include <algorithm>
#include <stdexcept>

#include "oracle.h"

namespace orc {

bool Vol::operator==(const Vol& o) const {
    return sx == o.sx && sy == o.sy && sz == o.sz && x0 == o.x0 && y0 == o.y0 && z0 == o.z0 && dx == o.dx && dy == o.dy && dz == o.dz;
}

namespace {
float unary(const std::string& op, float x) {
    if (op == "abs") return bf(fb(x) & 0x7FFFFFFFu);
    if (op == "square") return x * x;
    if (op == "cube") return x * x * x;
    if (op == "sqrt") return static_cast<float>(std::sqrt(static_cast<double>(x)));
    if (op == "half_negative") return x > 0.0f ? x : x * 0.5f;
    if (op == "quarter_negative") return x > 0.0f ? x : x * 0.25f;
    if (op == "reciprocal") return 1.0f / x;
    if (op == "negate") return -x;
    if (op == "squeeze") {
        const float c = fclamp(x, -1.0f, 1.0f);
        return c / 2.0f - c * c * c / 24.0f;
    }
    if (op == "sign") return (std::isnan(x) || x == 0.0f) ? x : (x > 0.0f ? 1.0f : -1.0f);
    throw std::runtime_error("unsupported unary op " + op);
}

float gradient(const Node& n, int c) {
    const int range = n.to - n.from;
    const float factor = (n.tv - n.fv) / static_cast<float>(range);
    if (n.tiling == "clamp_to_edge") {
        const int lo = std::min(n.from, n.to), hi = std::max(n.from, n.to);
        return n.fv + static_cast<float>(std::min(std::max(c, lo), hi) - n.from) * factor;
    }
    const int rel = wadd(c, -n.from);
    if (n.tiling == "repeat") return n.fv + static_cast<float>(floor_mod(rel, range)) * factor;
    const int tile = floor_div(rel, range);
    const int local = rel - tile * range;
    return n.fv + static_cast<float>((tile & 1) == 0 ? local : range - local) * factor;
}

int coord_of(const std::string& axis, int x, int y, int z) { return axis == "x" ? x : (axis == "y" ? y : z); }

// Mth.binarySearch(0, n, i -> in < loc[i]) - 1
int interval(const std::vector<float>& loc, float in) {
    int lo = 0, len = static_cast<int>(loc.size());
    while (len > 0) {
        const int half = len / 2, m = lo + half;
        if (in < loc[static_cast<size_t>(m)]) len = half;
        else { lo = m + 1; len -= half + 1; }
    }
    return lo - 1;
}

float extend(float in, const SplineN& s, float v, int i) {
    const float d = s.der[static_cast<size_t>(i)];
    return d == 0.0f ? v : v + d * (in - s.loc[static_cast<size_t>(i)]);
}

// CubicSpline.Multipoint between points i and i+1
float hermite(float in, const SplineN& s, int i, float y1, float y2) {
    const float x1 = s.loc[static_cast<size_t>(i)], x2 = s.loc[static_cast<size_t>(i) + 1];
    const float t = (in - x1) / (x2 - x1);
    const float d1 = s.der[static_cast<size_t>(i)], d2 = s.der[static_cast<size_t>(i) + 1];
    const float dx = x2 - x1, dy = y2 - y1;
    const float a = d1 * dx - dy, b = -d2 * dx + dy;
    return lerp(t, y1, y2) + t * (1.0f - t) * lerp(t, a, b);
}
}  // namespace

Density::Density(const Ir& ir, const Tables& t) : ir_(ir), t_(t) {}

void Density::fill_noise(const Stack& s, const Vol& v, double xz, double ys, float amp, float* acc) {
    int i = 0;
    for (int z = 0; z < v.sz; ++z)
        for (int x = 0; x < v.sx; ++x)
            for (int y = 0; y < v.sy; ++y, ++i) {
                float a = acc[i];
                for (int k = 0; k < s.count; ++k) {
                    const Layer& L = t_.layers[static_cast<size_t>(s.first + k)];
                    const Octave& o = t_.oct[static_cast<size_t>(L.octave)];
                    const int bx = v.x0 + x * v.dx, by = v.y0 + y * v.dy, bz = v.z0 + z * v.dz;
                    const float n = s.smeared ? noise_add_volume_smeared(o, bx, by, bz, xz * L.freq, ys * L.freq)
                                              : noise_add_volume(o, bx, by, bz, xz * L.freq, ys * L.freq);
                    a = a + (amp * L.amp) * n;
                }
                acc[i] = a;
            }
}

void Density::sample(int id, const Vol& v, float* out) {
    const Node& n = ir_.nodes[static_cast<size_t>(id)];
    const int N = v.n();
    auto buf = [&](int node) {
        std::vector<float> b(static_cast<size_t>(N));
        sample(node, v, b.data());
        return b;
    };
    auto each = [&](auto f) {   // f(index, x, y, z)
        int i = 0;
        for (int z = 0; z < v.sz; ++z)
            for (int x = 0; x < v.sx; ++x)
                for (int y = 0; y < v.sy; ++y, ++i) f(i, v.x0 + x * v.dx, v.y0 + y * v.dy, v.z0 + z * v.dz);
    };
    switch (n.t) {
        case T::Cache: {
            auto it = cache_.find(id);
            if (it == cache_.end() || !(it->second.first == v)) {
                std::vector<float> b(static_cast<size_t>(N), 0.0f);
                sample(n.in, v, b.data());
                cache_[id] = {v, std::move(b)};
                it = cache_.find(id);
            }
            std::copy(it->second.second.begin(), it->second.second.end(), out);
            return;
        }
        case T::Const: std::fill(out, out + N, n.v); return;
        case T::Simple: std::fill(out, out + N, n.kind == "blend_alpha" ? 1.0f : 0.0f); return;
        case T::Gradient: each([&](int i, int x, int y, int z) { out[i] = gradient(n, coord_of(n.axis, x, y, z)); }); return;
        case T::Noise: {
            const Stack& s = t_.noise.at(n.noise);
            if (n.zx && n.zy && n.zz) {
                std::fill(out, out + N, 0.0f);
                fill_noise(s, v, n.xz, n.ys, 1.0f, out);
                return;
            }
            const std::vector<float> bx = buf(n.sx), bz = buf(n.sz);
            std::vector<float> by;
            if (!n.zy) by = buf(n.sy);
            each([&](int i, int x, int y, int z) {
                const double px = static_cast<double>(x) * n.xz + static_cast<double>(bx[static_cast<size_t>(i)]);
                double py = static_cast<double>(y) * n.ys;
                if (!n.zy) py = py + static_cast<double>(by[static_cast<size_t>(i)]);
                const double pz = static_cast<double>(z) * n.xz + static_cast<double>(bz[static_cast<size_t>(i)]);
                out[i] = stack_get(t_, s, px, py, pz);
            });
            return;
        }
        case T::ShiftA:
        case T::Shift:
            std::fill(out, out + N, 0.0f);
            fill_noise(t_.noise.at(n.noise), v, 0.25, n.t == T::ShiftA ? 0.0 : 0.25, 1.0f, out);
            for (int i = 0; i < N; ++i) out[i] = out[i] * 4.0f;
            return;
        case T::ShiftB: {
            // the offset noise sampled with x from z, y from x, z = 0 (a transposed 2D volume)
            const Vol tv{v.sz, v.sx, 1, v.z0, v.x0, 0, v.dz, v.dx, 1};
            std::vector<float> tb(static_cast<size_t>(tv.n()), 0.0f);
            fill_noise(t_.noise.at(n.noise), tv, 0.25, 0.25, 4.0f, tb.data());
            each([&](int i, int x, int, int z) {
                out[i] = tb[static_cast<size_t>(tv.at((z - v.z0) / v.dz, (x - v.x0) / v.dx, 0))];
            });
            return;
        }
        case T::Blended: {
            const Stack* b = t_.blended.at(id);
            const double xzm = 684.412 * n.xz, ym = 684.412 * n.ys;
            std::vector<float> lo(static_cast<size_t>(N), 0.0f), hi(static_cast<size_t>(N), 0.0f);
            std::fill(out, out + N, 0.0f);
            fill_noise(b[2], v, xzm / n.xzf, ym / n.yf, 1.0f, out);
            fill_noise(b[0], v, xzm, ym, 1.0f, lo.data());
            fill_noise(b[1], v, xzm, ym, 1.0f, hi.data());
            for (int i = 0; i < N; ++i) {
                const float a = fclamp(out[i] + 0.5f, 0.0f, 1.0f);
                out[i] = a == 0.0f ? lo[static_cast<size_t>(i)] : (a == 1.0f ? hi[static_cast<size_t>(i)] : lerp(a, lo[static_cast<size_t>(i)], hi[static_cast<size_t>(i)]));
            }
            return;
        }
        case T::Binary: {
            const Node &L = ir_.nodes[static_cast<size_t>(n.l)], &R = ir_.nodes[static_cast<size_t>(n.r)];
            const bool lc = L.t == T::Const, rc = R.t == T::Const;
            if (n.op == "min" || n.op == "max") {
                const bool mn = n.op == "min";
                if (mn ? L.hi < R.lo : L.lo > R.hi) { sample(n.l, v, out); return; }
                if (mn ? R.hi < L.lo : R.lo > L.hi) { sample(n.r, v, out); return; }
                const int a = lc ? n.r : n.l, c = lc ? n.l : n.r;   // the buffered operand comes first
                sample(a, v, out);
                const std::vector<float> o2 = buf(c);
                for (int i = 0; i < N; ++i) {
                    const float q = o2[static_cast<size_t>(i)];
                    if (mn ? q < out[i] : q > out[i]) out[i] = q;
                }
                return;
            }
            sample(n.l, v, out);
            if (n.op == "div" && !lc && rc) {
                const float inv = 1.0f / R.v;
                for (int i = 0; i < N; ++i) out[i] = out[i] * inv;
                return;
            }
            const std::vector<float> o2 = buf(n.r);
            for (int i = 0; i < N; ++i) {
                const float q = o2[static_cast<size_t>(i)];
                if (n.op == "add") out[i] = out[i] + q;
                else if (n.op == "sub") out[i] = out[i] - q;
                else if (n.op == "mul") out[i] = out[i] * q;
                else if (n.op == "div") out[i] = out[i] / q;
                else throw std::runtime_error("binary op " + n.op);
            }
            return;
        }
        case T::Unary:
            sample(n.in, v, out);
            for (int i = 0; i < N; ++i) out[i] = unary(n.op, out[i]);
            return;
        case T::Clamp:
            sample(n.in, v, out);
            for (int i = 0; i < N; ++i) out[i] = fclamp(out[i], n.fmin, n.fmax);
            return;
        case T::Lerp: {
            sample(n.alpha, v, out);
            const std::vector<float> a = buf(n.first), b = buf(n.second);
            for (int i = 0; i < N; ++i) {
                const float t = out[i];
                out[i] = t == 0.0f ? a[static_cast<size_t>(i)] : (t == 1.0f ? b[static_cast<size_t>(i)] : lerp(t, a[static_cast<size_t>(i)], b[static_cast<size_t>(i)]));
            }
            return;
        }
        case T::RangeChoice: {
            const std::vector<float> in = buf(n.in), out_b = buf(n.when_out);
            sample(n.when_in, v, out);
            for (int i = 0; i < N; ++i) {
                const float c = in[static_cast<size_t>(i)];
                if (!(c >= n.fmin && c < n.fmax)) out[i] = out_b[static_cast<size_t>(i)];
            }
            return;
        }
        case T::IntervalSelect: {
            const std::vector<float> in = buf(n.in);
            std::vector<std::vector<float>> f;
            for (int fn : n.fns) f.push_back(buf(fn));
            for (int i = 0; i < N; ++i) {
                size_t k = 0;
                while (k < n.th.size() && !(in[static_cast<size_t>(i)] < n.th[k])) ++k;
                out[i] = f[k][static_cast<size_t>(i)];
            }
            return;
        }
        case T::Spline: {
            std::map<int, std::vector<float>> in;   // each coordinate sampled over the volume once
            for (int i = 0; i < N; ++i) out[i] = spline_vol(*n.spline, i, in, v);
            return;
        }
        case T::Interp: interpolated(n, v, out); return;
        case T::Slice: slice(n, v, out); return;
        case T::BlendDensity: sample(n.in, v, out); return;   // empty blender
        case T::Fts: {
            if (!(v.sy == 1 && v.y0 == 0)) {
                const Vol flat{v.sx, 1, v.sz, v.x0, 0, v.z0, v.dx, v.dy, v.dz};
                std::vector<float> fbuf(static_cast<size_t>(flat.n()));
                sample(id, flat, fbuf.data());
                each([&](int i, int x, int, int z) { out[i] = fbuf[static_cast<size_t>(flat.at((x - v.x0) / v.dx, 0, (z - v.z0) / v.dz))]; });
                return;
            }
            sample(n.upper, v, out);
            each([&](int i, int x, int, int z) { out[i] = surface_probe(n, x, z, out[i]); });
            return;
        }
    }
}

float Density::surface_probe(const Node& n, int x, int z, float upper) {
    const int top = floor_i(static_cast<double>(upper / static_cast<float>(n.cell_h))) * n.cell_h;
    if (top <= n.lower) return static_cast<float>(n.lower);
    for (int y = top; y >= n.lower; y -= n.cell_h)
        if (point(n.density, x, y, z) > 0.0f) return static_cast<float>(y);
    return static_cast<float>(n.lower);
}

float Density::spline_vol(const SplineN& s, int i, std::map<int, std::vector<float>>& in, const Vol& v) {
    if (s.is_const) return s.k;
    auto it = in.find(s.coord);
    if (it == in.end()) {
        std::vector<float> b(static_cast<size_t>(v.n()));
        sample(s.coord, v, b.data());
        it = in.emplace(s.coord, std::move(b)).first;
    }
    const float x = it->second[static_cast<size_t>(i)];
    const int k = interval(s.loc, x), last = static_cast<int>(s.loc.size()) - 1;
    if (k < 0) return extend(x, s, spline_vol(s.val.front(), i, in, v), 0);
    if (k == last) return extend(x, s, spline_vol(s.val[static_cast<size_t>(last)], i, in, v), last);
    const float y1 = spline_vol(s.val[static_cast<size_t>(k)], i, in, v);
    const float y2 = spline_vol(s.val[static_cast<size_t>(k) + 1], i, in, v);
    return hermite(x, s, k, y1, y2);
}

float Density::spline_pt(const SplineN& s, int x, int y, int z) {
    if (s.is_const) return s.k;
    const float c = point(s.coord, x, y, z);
    const int k = interval(s.loc, c), last = static_cast<int>(s.loc.size()) - 1;
    if (k < 0) return extend(c, s, spline_pt(s.val.front(), x, y, z), 0);
    if (k == last) return extend(c, s, spline_pt(s.val[static_cast<size_t>(last)], x, y, z), last);
    return hermite(c, s, k, spline_pt(s.val[static_cast<size_t>(k)], x, y, z), spline_pt(s.val[static_cast<size_t>(k) + 1], x, y, z));
}

// SliceFunction: the input on the plane axis = coord (X over Z compiles to one sampler)
void Density::slice(const Node& n, const Vol& v, float* out) {
    int ax = n.axis == "x", ay = n.axis == "y", az = n.axis == "z";
    int cx = n.coord, cz = n.coord;
    int inner = n.in;
    const Node& in = ir_.nodes[static_cast<size_t>(n.in)];
    if (in.t == T::Slice && ((ax && in.axis == "z") || (az && in.axis == "x"))) {
        cx = ax ? n.coord : in.coord;
        cz = az ? n.coord : in.coord;
        ax = az = 1;
        inner = in.in;
    }
    const Vol sv{ax ? 1 : v.sx, ay ? 1 : v.sy, az ? 1 : v.sz, ax ? cx : v.x0, ay ? n.coord : v.y0, az ? cz : v.z0, v.dx, v.dy, v.dz};
    if (sv == v) { sample(inner, v, out); return; }
    std::vector<float> b(static_cast<size_t>(sv.n()));
    sample(inner, sv, b.data());
    int i = 0;
    for (int z = 0; z < v.sz; ++z)
        for (int x = 0; x < v.sx; ++x)
            for (int y = 0; y < v.sy; ++y, ++i) out[i] = b[static_cast<size_t>(sv.at(ax ? 0 : x, ay ? 0 : y, az ? 0 : z))];
}

void Density::interpolated(const Node& n, const Vol& v, float* out) {
    const int c = n.cell_xz, cy = n.cell_y;
    const bool aligned = (v.dx == c || v.sx == 1) && (v.dy == cy || v.sy == 1) && (v.dz == c || v.sz == 1) && floor_mod(v.x0, c) == 0 &&
                         floor_mod(v.y0, cy) == 0 && floor_mod(v.z0, c) == 0;
    if (aligned) { sample(n.in, v, out); return; }
    if (v.dx == 1 && v.dy == 1 && v.dz == 1) { interp_blocks(n, v, out); return; }
    const Vol bv{v.sx * v.dx, v.sy * v.dy, v.sz * v.dz, v.x0, v.y0, v.z0, 1, 1, 1};
    std::vector<float> b(static_cast<size_t>(bv.n()));
    interp_blocks(n, bv, b.data());
    int i = 0;
    for (int z = 0; z < v.sz; ++z)
        for (int x = 0; x < v.sx; ++x)
            for (int y = 0; y < v.sy; ++y, ++i) out[i] = b[static_cast<size_t>(bv.at(x * v.dx, y * v.dy, z * v.dz))];
}

// InterpolatedFunction over a block-step volume: corner values on the cell lattice, then per
// cell z, x lerps and y as a running sum of steps.
void Density::interp_blocks(const Node& n, const Vol& v, float* out) {
    const int c = n.cell_xz, cy = n.cell_y;
    const float ixz = 1.0f / static_cast<float>(c), iy = 1.0f / static_cast<float>(cy);
    const int mx = v.x0 + v.sx - 1, my = v.y0 + v.sy - 1, mz = v.z0 + v.sz - 1;
    const int cx0 = floor_div(v.x0, c), cy0 = floor_div(v.y0, cy), cz0 = floor_div(v.z0, c);
    const int nx = floor_div(mx, c) - cx0 + 1, ny = floor_div(my, cy) - cy0 + 1, nz = floor_div(mz, c) - cz0 + 1;
    const Vol cv{floor_mod(mx, c) == 0 ? nx : nx + 1, floor_mod(my, cy) == 0 ? ny : ny + 1, floor_mod(mz, c) == 0 ? nz : nz + 1,
                 cx0 * c, cy0 * cy, cz0 * c, c, cy, c};
    std::vector<float> k(static_cast<size_t>(cv.n()));
    sample(n.in, cv, k.data());
    auto K = [&](int x, int y, int z) { return k[static_cast<size_t>(cv.at(std::min(x, cv.sx - 1), std::min(y, cv.sy - 1), std::min(z, cv.sz - 1)))]; };
    for (int gz = 0; gz < nz; ++gz)
        for (int gx = 0; gx < nx; ++gx)
            for (int gy = 0; gy < ny; ++gy) {
                const float v000 = K(gx, gy, gz), v100 = K(gx + 1, gy, gz), v001 = K(gx, gy, gz + 1), v101 = K(gx + 1, gy, gz + 1);
                const float v010 = K(gx, gy + 1, gz), v110 = K(gx + 1, gy + 1, gz), v011 = K(gx, gy + 1, gz + 1), v111 = K(gx + 1, gy + 1, gz + 1);
                const int ox = (cx0 + gx) * c - v.x0, oy = (cy0 + gy) * cy - v.y0, oz = (cz0 + gz) * c - v.z0;
                const int xa = std::max(0, -ox), ya = std::max(0, -oy), za = std::max(0, -oz);
                const int xb = std::min(c, v.sx - ox), yb = std::min(cy, v.sy - oy), zb = std::min(c, v.sz - oz);
                for (int z = za; z < zb; ++z) {
                    const float tz = static_cast<float>(z) * ixz;
                    const float a00 = lerp(tz, v000, v001), a01 = lerp(tz, v010, v011), a10 = lerp(tz, v100, v101), a11 = lerp(tz, v110, v111);
                    for (int x = xa; x < xb; ++x) {
                        const float tx = static_cast<float>(x) * ixz;
                        const float lo = lerp(tx, a00, a10), hi = lerp(tx, a01, a11);
                        const float step = (hi - lo) * iy;
                        float val = lo + step * static_cast<float>(ya);
                        for (int y = ya; y < yb; ++y) {
                            out[v.at(ox + x, oy + y, oz + z)] = val;
                            val += step;
                        }
                    }
                }
            }
}

float Density::point(int id, int x, int y, int z) {
    const Node& n = ir_.nodes[static_cast<size_t>(id)];
    switch (n.t) {
        case T::Cache:
        case T::BlendDensity: return point(n.in, x, y, z);
        case T::Const: return n.v;
        case T::Simple: return n.kind == "blend_alpha" ? 1.0f : 0.0f;
        case T::Gradient: return gradient(n, coord_of(n.axis, x, y, z));
        case T::Noise: {
            const Stack& s = t_.noise.at(n.noise);
            if (n.zx && n.zy && n.zz) return stack_get(t_, s, x * n.xz, y * n.ys, z * n.xz);
            const double px = x * n.xz + static_cast<double>(point(n.sx, x, y, z));
            double py = y * n.ys;
            if (!n.zy) py = py + static_cast<double>(point(n.sy, x, y, z));
            const double pz = z * n.xz + static_cast<double>(point(n.sz, x, y, z));
            return stack_get(t_, s, px, py, pz);
        }
        case T::ShiftA: return stack_get(t_, t_.noise.at(n.noise), x * 0.25, y * 0.0, z * 0.25) * 4.0f;
        case T::ShiftB: return stack_get(t_, t_.noise.at(n.noise), z * 0.25, x * 0.25, 0.0) * 4.0f;
        case T::Shift: return stack_get(t_, t_.noise.at(n.noise), x * 0.25, y * 0.25, z * 0.25) * 4.0f;
        case T::Blended: {
            const Stack* b = t_.blended.at(id);
            const double xzm = 684.412 * n.xz, ym = 684.412 * n.ys, mxz = xzm / n.xzf, my = ym / n.yf;
            const float a = fclamp(stack_get(t_, b[2], x * mxz, y * my, z * mxz) + 0.5f, 0.0f, 1.0f);
            if (a == 0.0f) return stack_get(t_, b[0], x * xzm, y * ym, z * xzm);
            if (a == 1.0f) return stack_get(t_, b[1], x * xzm, y * ym, z * xzm);
            const float lo = stack_get(t_, b[0], x * xzm, y * ym, z * xzm), hi = stack_get(t_, b[1], x * xzm, y * ym, z * xzm);
            return lerp(a, lo, hi);
        }
        case T::Binary: {
            const Node &L = ir_.nodes[static_cast<size_t>(n.l)], &R = ir_.nodes[static_cast<size_t>(n.r)];
            const bool lc = L.t == T::Const, rc = R.t == T::Const;
            if (n.op == "add") return point(n.l, x, y, z) + point(n.r, x, y, z);
            if (n.op == "sub") return point(n.l, x, y, z) - point(n.r, x, y, z);
            if (n.op == "mul") {
                if (lc || rc) return point(n.l, x, y, z) * point(n.r, x, y, z);
                const float a = point(n.l, x, y, z);
                return a == 0.0f ? 0.0f : a * point(n.r, x, y, z);
            }
            if (n.op == "div") {
                if (lc) return L.v / point(n.r, x, y, z);
                if (rc) return point(n.l, x, y, z) * (1.0f / R.v);
                const float a = point(n.l, x, y, z);
                return a == 0.0f ? 0.0f : a / point(n.r, x, y, z);
            }
            const bool mn = n.op == "min";
            if (mn ? L.hi < R.lo : L.lo > R.hi) return point(n.l, x, y, z);
            if (mn ? R.hi < L.lo : R.lo > L.hi) return point(n.r, x, y, z);
            if (lc || rc) {
                const float s = point(lc ? n.r : n.l, x, y, z), c = lc ? L.v : R.v;
                return mn ? java_min(s, c) : java_max(s, c);
            }
            const float a = point(n.l, x, y, z);
            if (mn) return a <= R.lo ? a : java_min(a, point(n.r, x, y, z));
            return a >= R.hi ? a : java_max(a, point(n.r, x, y, z));
        }
        case T::Unary: return unary(n.op, point(n.in, x, y, z));
        case T::Clamp: return fclamp(point(n.in, x, y, z), n.fmin, n.fmax);
        case T::Lerp: {
            const float a = point(n.alpha, x, y, z);
            if (a == 0.0f) return point(n.first, x, y, z);
            if (a == 1.0f) return point(n.second, x, y, z);
            const float p = point(n.first, x, y, z), q = point(n.second, x, y, z);
            return lerp(a, p, q);
        }
        case T::RangeChoice: {
            const float c = point(n.in, x, y, z);
            return (c >= n.fmin && c < n.fmax) ? point(n.when_in, x, y, z) : point(n.when_out, x, y, z);
        }
        case T::IntervalSelect: {
            const float c = point(n.in, x, y, z);
            size_t k = 0;
            while (k < n.th.size() && !(c < n.th[k])) ++k;
            return point(n.fns[k], x, y, z);
        }
        case T::Spline: return spline_pt(*n.spline, x, y, z);
        case T::Interp: {
            const int xi = floor_mod(x, n.cell_xz), yi = floor_mod(y, n.cell_y), zi = floor_mod(z, n.cell_xz);
            if (xi == 0 && yi == 0 && zi == 0) return point(n.in, x, y, z);
            const Vol cv{2, 2, 2, x - xi, y - yi, z - zi, n.cell_xz, n.cell_y, n.cell_xz};
            float k[8];
            sample(n.in, cv, k);
            return lerp3(static_cast<float>(xi) / static_cast<float>(n.cell_xz), static_cast<float>(yi) / static_cast<float>(n.cell_y),
                         static_cast<float>(zi) / static_cast<float>(n.cell_xz), k[cv.at(0, 0, 0)], k[cv.at(1, 0, 0)], k[cv.at(0, 1, 0)], k[cv.at(1, 1, 0)],
                         k[cv.at(0, 0, 1)], k[cv.at(1, 0, 1)], k[cv.at(0, 1, 1)], k[cv.at(1, 1, 1)]);
        }
        case T::Slice:
            return point(n.in, n.axis == "x" ? n.coord : x, n.axis == "y" ? n.coord : y, n.axis == "z" ? n.coord : z);
        case T::Fts: return surface_probe(n, x, z, point(n.upper, x, 0, z));
    }
    throw std::runtime_error("unhandled node");
}

}  // namespace orc
