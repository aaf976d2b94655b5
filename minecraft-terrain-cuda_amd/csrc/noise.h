// Improved-Perlin gradient noise as the TERRAIN step samples it, host and device.
//
// Two sampling paths exist and must not be mixed (they round differently):
//   point  (PerlinNoise.get): coordinates (block * scale) * frequency, gradient dot
//          (gx*x + gy*y) + gz*z, trilinear lerp at the point;
//   volume (PerlinNoise.addToVolume): coordinates block * (scale * frequency), dot
//          (gx*x + gz*z) + gy*y with the xz part shared along y.
// Smeared octaves (BlendedNoise) shift the y fraction by SmearedPerlinNoise.computeFudgeY.
// Reference semantics: cpu/include/mccuda/noise_eval.h:21-304.
#pragma once

#include <stddef.h>

#include "jmath.h"

namespace mct {

// One gradient-noise octave of one seed: random origin and permutation p. The permutation is
// stored as entry pairs q[i] = p[i & 255] | p[(i + 1) & 255] << 8 for i in [0, 511]: the
// lattice hashes index it without masking (with the cell coordinates reduced mod 256 once,
// every lookup index is at most 511, as in a doubled table), and each of their lookups goes
// in adjacent pairs (p[k], p[k + 1]), one 16-bit load each. A hash sample then takes 6 or 7
// loads instead of 12 or 14; the loads are what bounds the noise kernels (their vector
// memory data path, not their arithmetic).
struct Octave {
    double ox, oy, oz;
    uint16_t q[512];
};

// GRADIENT[h & 15] components as signed 2-bit fields (two's complement: -1 = 0b11), entry h
// at bits 2h..2h+1, so one sign-extending bit-field extract gives the component:
//   gx: 1,-1,1,-1,1,-1,1,-1,0,0,0,0,1,0,-1,0
//   gy: 1,1,-1,-1,0,0,0,0,1,-1,1,-1,1,-1,1,-1
//   gz: 0,0,0,0,1,1,-1,-1,1,1,-1,-1,0,1,0,-1
constexpr uint32_t kGX = 0x3100DDDDu, kGY = 0xDDDD00F5u, kGZ = 0xC4F5F500u;
MCT_HD MCT_FI float gcomp(uint32_t packed, int h) {
#if defined(__HIP_DEVICE_COMPILE__)
    // v_bfe_i32 written out: the compiler's own sign-extending extract, followed by the int to
    // float conversion, was compiled to an unsigned conversion (-1 became 4294967295.0). The
    // instruction reads the offset's low 5 bits only, so 2h needs no mask: they are 2 (h & 15).
    int g;
    asm("v_bfe_i32 %0, %1, %2, 2" : "=v"(g) : "s"(packed), "v"(h << 1));
    return static_cast<float>(g);
#else
    return static_cast<float>(static_cast<int32_t>(packed << (30 - ((h & 15) << 1))) >> 30);
#endif
}

// Math.nextDown(2^24): GradientNoise.wrap leaves |x| below it untouched.
constexpr double kWrapLimit = 0x1.fffffffffffffp+23;   // bits 0x416FFFFFFFFFFFFF

// GradientNoise.wrap: x - floor(x / 2^25 + 1/2) * 2^25 outside [-limit, limit). On the device
// the fold is skipped for the whole wave when no lane needs it (one vote).
MCT_HD MCT_FI double wrap(double x) {
    const bool in = x >= -kWrapLimit && x < kWrapLimit;
#if defined(__HIP_DEVICE_COMPILE__)
    if (__builtin_amdgcn_ballot_w64(!in) == 0) return x;
#endif
    if (in) return x;
    double q = x / 33554432.0;
    q = q + 0.5;
    q = floor(q);
    q = q * 33554432.0;
    return x - q;
}

// Floor of a wrapped lattice coordinate (|v| < 2^24 + 256 after wrap plus an offset in
// [0, 256), never NaN): Mth.floor without the saturation and NaN cases, which cannot occur.
MCT_HD MCT_FI int32_t lattice_floor(double v) { return static_cast<int32_t>(floor(v)); }

// Permutation lookups of one octave, index in [0, 511]: P(i) = p[i & 255], P.pair(i) =
// P(i) | P(i + 1) << 8. PermRef reads through the octave pointer.
struct PermRef {
    const uint16_t* q;
    MCT_HD MCT_FI int operator()(int32_t i) const { return q[i] & 0xFF; }
    MCT_HD MCT_FI uint32_t pair(int32_t i) const { return q[i]; }
};

#if defined(__HIP_DEVICE_COMPILE__)
// The first active lane's octave-table pointer (wave-uniform, in SGPRs).
__device__ MCT_FI const Octave* wave_first(const Octave* p) {
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(v))));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(static_cast<uint32_t>(v >> 32))));
    return reinterpret_cast<const Octave*>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Buffer resource over a wave-uniform table of n octaves (gfx9 raw buffer, stride 0,
// bounds-checked to the table).
__device__ MCT_FI __amdgpu_buffer_rsrc_t octave_rsrc(const Octave* table, int n) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<Octave*>(table), 0, n * static_cast<int>(sizeof(Octave)), 0x00020000);
}

// PermBuf: lookups of octave k of a wave-uniform table as raw buffer 16-bit loads, the
// octave's byte offset as scalar offset and the pair index as 32-bit vector offset (no
// per-lane 64-bit address arithmetic).
struct PermBuf {
    __amdgpu_buffer_rsrc_t r;
    int off;
    __device__ MCT_FI PermBuf(__amdgpu_buffer_rsrc_t rsrc, int k) : r(rsrc), off(k * static_cast<int>(sizeof(Octave)) + static_cast<int>(offsetof(Octave, q))) {}
    __device__ MCT_FI uint32_t pair(int32_t i) const { return __builtin_amdgcn_raw_buffer_load_b16(r, i * 2, off, 0); }
    __device__ MCT_FI int operator()(int32_t i) const { return static_cast<int>(pair(i) & 0xFF); }
};
#endif

// Lattice hashes of the 8 cell corners, h[xyz] with bit 4 = x, 2 = y, 1 = z. Only bits 0-3
// of a hash are meaningful (the gradient index h & 15); higher bits may be set.
struct Hash8 {
    int h000, h100, h010, h110, h001, h101, h011, h111;
};

// p[(a + f) & 255] for the Java int sums a + f of the reference equals P(a + (f & 255)) (the
// pair table covers indexes up to 511, a <= 255), so each coordinate is reduced once and no
// lookup is masked. The lookups go in adjacent pairs (p[k], p[k + 1]), one load each.
template <class Perm>
MCT_HD MCT_FI Hash8 hash8_pairs(const Perm& P, int a, int b, int Y, int Z) {
    const uint32_t A = P.pair(a + Y), B = P.pair(b + Y);   // a0 | a1 << 8, b0 | b1 << 8
    const uint32_t H0 = P.pair(static_cast<int>(A & 0xFF) + Z), H1 = P.pair(static_cast<int>(B & 0xFF) + Z);
    const uint32_t H2 = P.pair(static_cast<int>(A >> 8) + Z), H3 = P.pair(static_cast<int>(B >> 8) + Z);
    // the gradient picks read bits 0-3 of a hash only, so the z = 0 corners keep the pair's
    // upper byte instead of a mask
    Hash8 h;
    h.h000 = static_cast<int>(H0);
    h.h001 = static_cast<int>(H0 >> 8);
    h.h100 = static_cast<int>(H1);
    h.h101 = static_cast<int>(H1 >> 8);
    h.h010 = static_cast<int>(H2);
    h.h011 = static_cast<int>(H2 >> 8);
    h.h110 = static_cast<int>(H3);
    h.h111 = static_cast<int>(H3 >> 8);
    return h;
}

template <class Perm>
MCT_HD MCT_FI Hash8 hash8(const Perm& P, int32_t fx, int32_t fy, int32_t fz) {
    const uint32_t ab = P.pair(fx & 0xFF);
    return hash8_pairs(P, static_cast<int>(ab & 0xFF), static_cast<int>(ab >> 8), fy & 0xFF, fz & 0xFF);
}

// hash8 from the first-level entries a = p[x & 255], b = p[(x & 255) + 1].
template <class Perm>
MCT_HD MCT_FI Hash8 hash8_ab(const Perm& P, int a, int b, int32_t fy, int32_t fz) {
    return hash8_pairs(P, a, b, fy & 0xFF, fz & 0xFF);
}

// g * v + c for a gradient component g in {-1, 0, 1}. The product is exact (v, -v or a zero
// with the product's sign), so RN(g * v + c) under one fused multiply-add equals the unfused
// RN(RN(g * v) + c), the sign of a zero sum included: the device fuses, the host does not,
// and both give the reference's sums bit for bit.
MCT_HD MCT_FI float gmad(uint32_t packed, int h, float v, float c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_fmaf(gcomp(packed, h), v, c);
#else
    return gcomp(packed, h) * v + c;
#endif
}

// Gradient dot for the point path: (gx*x + gy*y) + gz*z.
MCT_HD MCT_FI float dot3(int h, float x, float y, float z) {
    return gmad(kGZ, h, z, gmad(kGY, h, y, gcomp(kGX, h) * x));
}

// PerlinNoise.sampleAndLerp (point path); ty is the y fraction the smoothstep uses.
template <class Perm>
MCT_HD MCT_FI float sample_lerp(const Perm& P, int32_t fx, int32_t fy, int32_t fz, float rx, float ry, float rz, float ty) {
    const Hash8 h = hash8(P, fx, fy, fz);
    const float rx1 = rx - 1.0f, ry1 = ry - 1.0f, rz1 = rz - 1.0f;
    const float d000 = dot3(h.h000, rx, ry, rz);
    const float d100 = dot3(h.h100, rx1, ry, rz);
    const float d010 = dot3(h.h010, rx, ry1, rz);
    const float d110 = dot3(h.h110, rx1, ry1, rz);
    const float d001 = dot3(h.h001, rx, ry, rz1);
    const float d101 = dot3(h.h101, rx1, ry, rz1);
    const float d011 = dot3(h.h011, rx, ry1, rz1);
    const float d111 = dot3(h.h111, rx1, ry1, rz1);
    return lerp3f(smooth(rx), smooth(ty), smooth(rz), d000, d100, d010, d110, d001, d101, d011, d111);
}

// PerlinNoise.get(x, y, z); P: the octave's permutation lookups.
template <class Perm>
MCT_HD MCT_FI float perlin_point(const Octave& o, const Perm& P, double px, double py, double pz) {
    const double x = wrap(px) + o.ox, y = wrap(py) + o.oy, z = wrap(pz) + o.oz;
    const int32_t fx = lattice_floor(x), fy = lattice_floor(y), fz = lattice_floor(z);
    const float rx = static_cast<float>(x - static_cast<double>(fx));
    const float ry = static_cast<float>(y - static_cast<double>(fy));
    const float rz = static_cast<float>(z - static_cast<double>(fz));
    return sample_lerp(P, fx, fy, fz, rx, ry, rz, ry);
}

MCT_HD MCT_FI float perlin_point(const Octave& o, double px, double py, double pz) { return perlin_point(o, PermRef{o.q}, px, py, pz); }

// SmearedPerlinNoise.computeFudgeY
MCT_HD MCT_FI double fudge_y(double original_y, double rel_y, double smear) {
    const double limit = (original_y >= 0.0 && original_y < rel_y) ? original_y : rel_y;
    double q = limit / smear;
    q = q + static_cast<double>(1.0E-7f);
    return static_cast<double>(mfloor(q)) * smear;
}

// SmearedPerlinNoise.get(x, y, z)
template <class Perm>
MCT_HD MCT_FI float smeared_point(const Octave& o, const Perm& P, double smear, double px, double py, double pz) {
    const double x = wrap(px) + o.ox, y = wrap(py) + o.oy, z = wrap(pz) + o.oz;
    const int32_t fx = lattice_floor(x), fy = lattice_floor(y), fz = lattice_floor(z);
    const float rx = static_cast<float>(x - static_cast<double>(fx));
    const double rel_y = y - static_cast<double>(fy);
    const float rz = static_cast<float>(z - static_cast<double>(fz));
    const float fudged = static_cast<float>(rel_y - fudge_y(py, rel_y, smear));
    return sample_lerp(P, fx, fy, fz, rx, fudged, rz, static_cast<float>(rel_y));
}

MCT_HD MCT_FI float smeared_point(const Octave& o, double smear, double px, double py, double pz) {
    return smeared_point(o, PermRef{o.q}, smear, px, py, pz);
}

// Coordinate setup of one axis of the volume path, the only FP64 work of a sample:
// lattice cell, fraction handed to the gradients, smoothstep weight. It depends on one
// block coordinate and one octave, so it can be tabulated per distinct coordinate.
struct Axis {
    int32_t f;
    float r;
    float a;
};

// The x and z records of one octave sample site of a block column. X.f is not the x lattice
// cell but the two first-level permutation entries of it, p[x & 255] | p[(x & 255) + 1] << 16
// (perm_x): the hash of every sample of the column starts from them.
struct AxisXZ {
    Axis X, Z;
};

// x and z (and y of plain octaves): v = wrap(b * s) + offset
MCT_HD MCT_FI Axis axis_plain(int32_t b, double s, double offset) {
    const double v = wrap(static_cast<double>(b) * s) + offset;
    const int32_t f = lattice_floor(v);
    const float r = static_cast<float>(v - static_cast<double>(f));
    return {f, r, smooth(r)};
}

// y of smeared octaves: the smoothstep uses the true fraction, the gradients the fudged one.
MCT_HD MCT_FI Axis axis_smeared(int32_t b, double s, double offset, double smear) {
    const double orig = static_cast<double>(b) * s;
    const double v = wrap(orig) + offset;
    const int32_t f = lattice_floor(v);
    const double rel = v - static_cast<double>(f);
    const float a = smooth(static_cast<float>(rel));
    const float r = static_cast<float>(rel - fudge_y(orig, rel, smear));
    return {f, r, a};
}

// Volume-path sample from the hashes and the three axis records: xz dots per corner, plus gy * ry.
MCT_HD MCT_FI float volume_corners(const Hash8& h, const Axis& X, const Axis& Y, const Axis& Z) {
    const float rx = X.r, rz = Z.r, rx1 = rx - 1.0f, rz1 = rz - 1.0f;
    const float ry = Y.r, ry1 = ry - 1.0f;
#define MCT_XZY(hh, xx, yy, zz) gmad(kGY, hh, yy, gmad(kGZ, hh, zz, gcomp(kGX, hh) * (xx)))
    const float c000 = MCT_XZY(h.h000, rx, ry, rz);
    const float c100 = MCT_XZY(h.h100, rx1, ry, rz);
    const float c010 = MCT_XZY(h.h010, rx, ry1, rz);
    const float c110 = MCT_XZY(h.h110, rx1, ry1, rz);
    const float c001 = MCT_XZY(h.h001, rx, ry, rz1);
    const float c101 = MCT_XZY(h.h101, rx1, ry, rz1);
    const float c011 = MCT_XZY(h.h011, rx, ry1, rz1);
    const float c111 = MCT_XZY(h.h111, rx1, ry1, rz1);
#undef MCT_XZY
    return lerp3f(X.a, Y.a, Z.a, c000, c100, c010, c110, c001, c101, c011, c111);
}

#if defined(__HIPCC__)
// GRADIENT[h & 15] as the float bits of (gx, gy, gz) and a zero word in LDS, for the kernels
// that fill it first (grad_table_fill): one 16-byte LDS read per corner instead of three
// bit-field extracts and three conversions on the VALU. The 16 entries are few enough that the
// lanes of a wave mostly share them (broadcast reads). The zero word is OR-ed into gy's bits,
// which keeps the read at 16 bytes (a 12-byte LDS read is several times slower).
__shared__ int4 g_grad[16];
__device__ MCT_FI void grad_table_fill(int t) {
    if (t < 16) g_grad[t] = make_int4(static_cast<int>(f2u(gcomp(kGX, t))), static_cast<int>(f2u(gcomp(kGY, t))), static_cast<int>(f2u(gcomp(kGZ, t))), 0);
}

// volume_corners with the gradients read from g_grad: the same products and fused sums.
__device__ MCT_FI float volume_corners_tab(const Hash8& h, const Axis& X, const Axis& Y, const Axis& Z) {
    const float rx = X.r, rz = Z.r, rx1 = rx - 1.0f, rz1 = rz - 1.0f;
    const float ry = Y.r, ry1 = ry - 1.0f;
    const auto dot = [](int hh, float xx, float yy, float zz) {
        const int4 g = g_grad[hh & 15];
        const float gx = u2f(static_cast<uint32_t>(g.x)), gy = u2f(static_cast<uint32_t>(g.y | g.w)), gz = u2f(static_cast<uint32_t>(g.z));
        return __builtin_fmaf(gy, yy, __builtin_fmaf(gz, zz, gx * xx));
    };
    const float c000 = dot(h.h000, rx, ry, rz);
    const float c100 = dot(h.h100, rx1, ry, rz);
    const float c010 = dot(h.h010, rx, ry1, rz);
    const float c110 = dot(h.h110, rx1, ry1, rz);
    const float c001 = dot(h.h001, rx, ry, rz1);
    const float c101 = dot(h.h101, rx1, ry, rz1);
    const float c011 = dot(h.h011, rx, ry1, rz1);
    const float c111 = dot(h.h111, rx1, ry1, rz1);
    return lerp3f(X.a, Y.a, Z.a, c000, c100, c010, c110, c001, c101, c011, c111);
}
#endif

template <class Perm>
MCT_HD MCT_FI float volume_sample(const Perm& P, const Axis& X, const Axis& Y, const Axis& Z) {
    return volume_corners(hash8(P, X.f, Y.f, Z.f), X, Y, Z);
}

// The same from a column table record (AxisXZ, X.f holding perm_x).
template <class Perm>
MCT_HD MCT_FI float volume_sample_xz(const Perm& P, const AxisXZ& A, const Axis& Y) {
    return volume_corners(hash8_ab(P, A.X.f & 0xFFFF, A.X.f >> 16, Y.f, A.Z.f), A.X, Y, A.Z);
}

// volume_sample_xz for the kernels that hold the gradient table in LDS (on the host, the same
// function as volume_sample_xz).
template <class Perm>
MCT_HD MCT_FI float volume_sample_xz_tab(const Perm& P, const AxisXZ& A, const Axis& Y) {
#if defined(__HIP_DEVICE_COMPILE__)
    return volume_corners_tab(hash8_ab(P, A.X.f & 0xFFFF, A.X.f >> 16, Y.f, A.Z.f), A.X, Y, A.Z);
#else
    return volume_sample_xz(P, A, Y);
#endif
}

// PerlinNoise.addToVolume, one sample (before the amplitude multiply).
template <class Perm>
MCT_HD MCT_FI float perlin_volume(const Octave& o, const Perm& P, int32_t bx, int32_t by, int32_t bz, double xs, double ys) {
    const Axis Z = axis_plain(bz, xs, o.oz);
    const Axis X = axis_plain(bx, xs, o.ox);
    const Axis Y = axis_plain(by, ys, o.oy);
    return volume_sample(P, X, Y, Z);
}

MCT_HD MCT_FI float perlin_volume(const Octave& o, int32_t bx, int32_t by, int32_t bz, double xs, double ys) {
    return perlin_volume(o, PermRef{o.q}, bx, by, bz, xs, ys);
}

// SmearedPerlinNoise.addToVolume, one sample.
template <class Perm>
MCT_HD MCT_FI float smeared_volume(const Octave& o, const Perm& P, double smear, int32_t bx, int32_t by, int32_t bz, double xs, double ys) {
    const Axis Z = axis_plain(bz, xs, o.oz);
    const Axis X = axis_plain(bx, xs, o.ox);
    const Axis Y = axis_smeared(by, ys, o.oy, smear);
    return volume_sample(P, X, Y, Z);
}

MCT_HD MCT_FI float smeared_volume(const Octave& o, double smear, int32_t bx, int32_t by, int32_t bz, double xs, double ys) {
    return smeared_volume(o, PermRef{o.q}, smear, bx, by, bz, xs, ys);
}

}  // namespace mct
