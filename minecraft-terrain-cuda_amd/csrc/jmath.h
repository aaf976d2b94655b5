// Java-language arithmetic semantics used by the TERRAIN step, for host and gfx950 device
// code. Every function reproduces the JLS / java.lang.Math / Mth operation it names bit for
// bit, given IEEE-754 round-to-nearest, no FMA contraction (-ffp-contract=off), no
// flush-to-zero and correctly rounded division/sqrt.
// Reference semantics: cpu/include/mccuda/javamath.h:61-181, cpu/include/mccuda/df_ops.h:14-80.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MCT_HD __host__ __device__
#define MCT_FI __forceinline__
#define MCT_NI __noinline__
#else
#define MCT_HD
#define MCT_FI inline __attribute__((always_inline))
#define MCT_NI __attribute__((noinline))
#endif

// Closures the generated evaluators define for memoised values: always inlined.
#define MCT_CLOSURE __attribute__((always_inline))

// Operation-count hooks. The host build compiled with MCT_COUNT_OPS adds each hook's counts
// (FP64, FP32, INT32 operations, memory reads) to a per-thread tally each time it runs; every
// other build, the device's included, compiles them to nothing.
#if defined(MCT_COUNT_OPS) && !defined(__HIP_DEVICE_COMPILE__)
namespace mct {
struct OpTally { unsigned long long f64, f32, i32, mem; };
inline thread_local OpTally g_tally{};
}  // namespace mct
#define MCT_OPS(a, b, c, d) ((void)(mct::g_tally.f64 += (a), mct::g_tally.f32 += (b), mct::g_tally.i32 += (c), mct::g_tally.mem += (d)))
#else
#define MCT_OPS(a, b, c, d) ((void)0)
#endif

// Pointers into LDS (workgroup memory) on the device, so that loads through them are ds_read
// even across out-of-line calls; plain pointers on the host.
#if defined(__HIP_DEVICE_COMPILE__)
#define MCT_LDS __attribute__((address_space(3)))
#else
#define MCT_LDS
#endif

namespace mct {

MCT_HD MCT_FI uint32_t f2u(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
MCT_HD MCT_FI float u2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }
MCT_HD MCT_FI uint64_t d2u(double d) { uint64_t u; __builtin_memcpy(&u, &d, 8); return u; }
MCT_HD MCT_FI double u2d(uint64_t u) { double d; __builtin_memcpy(&d, &u, 8); return d; }

// (int) v for a double (JLS 5.1.3): NaN -> 0, saturating, otherwise toward zero.
MCT_HD MCT_FI int32_t d2i(double v) {
    if (!(v == v)) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return -2147483647 - 1;
    return static_cast<int32_t>(v);
}

// Mth.floor(double) = (int) Math.floor(v). floor() is exact; the cast saturates.
MCT_HD MCT_FI int32_t mfloor(double v) { return d2i(floor(v)); }

// Java int + and * wrap modulo 2^32.
MCT_HD MCT_FI int32_t iadd(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) + static_cast<uint32_t>(b)); }
MCT_HD MCT_FI int32_t imul(int32_t a, int32_t b) { return static_cast<int32_t>(static_cast<uint32_t>(a) * static_cast<uint32_t>(b)); }

// Math.floorDiv / floorMod (int).
MCT_HD MCT_FI int32_t fdiv(int32_t a, int32_t b) {
    int32_t q = a / b;
    return ((a % b) != 0 && ((a ^ b) < 0)) ? q - 1 : q;
}
MCT_HD MCT_FI int32_t fmod_(int32_t a, int32_t b) { return a - fdiv(a, b) * b; }

// Math.min / Math.max (float): NaN wins, -0.0 < +0.0. Written as selects, so that the device
// code has no branch: a NaN a is kept by the last select, a NaN b falls through the compare
// (a <= NaN is false) and is picked, and of two zeros the one Java returns is picked.
MCT_HD MCT_FI float jminf(float a, float b) {
    float r = a <= b ? a : b;
    r = (a == 0.0f && b == 0.0f && (f2u(b) >> 31)) ? b : r;
    return a != a ? a : r;
}
MCT_HD MCT_FI float jmaxf(float a, float b) {
    float r = a >= b ? a : b;
    r = (a == 0.0f && b == 0.0f && (f2u(a) >> 31)) ? b : r;
    return a != a ? a : r;
}
MCT_HD MCT_FI double jmind(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && (d2u(b) >> 63)) return b;
    return a <= b ? a : b;
}
MCT_HD MCT_FI double jmaxd(double a, double b) {
    if (a != a) return a;
    if (a == 0.0 && b == 0.0 && (d2u(a) >> 63)) return b;
    return a >= b ? a : b;
}
MCT_HD MCT_FI int32_t jmini(int32_t a, int32_t b) { return a <= b ? a : b; }
MCT_HD MCT_FI int32_t jmaxi(int32_t a, int32_t b) { return a >= b ? a : b; }

// Math.abs(float) clears the sign bit (also of NaN and -0.0).
MCT_HD MCT_FI float jabsf(float a) { return u2f(f2u(a) & 0x7FFFFFFFu); }

// Mth.clamp(float|double|int): v < lo ? lo : Math.min(v, hi)
// For v not below lo, Math.min(v, hi) is hi exactly when v > hi, when hi is NaN and v is not,
// or when both are zeros and hi is -0.0 (a NaN v is kept); as selects.
MCT_HD MCT_FI float clampf(float v, float lo, float hi) {
    const bool to_hi = v > hi || (hi != hi && v == v) || (v == 0.0f && hi == 0.0f && (f2u(hi) >> 31));
    return v < lo ? lo : (to_hi ? hi : v);
}
MCT_HD MCT_FI double clampd(double v, double lo, double hi) { return v < lo ? lo : jmind(v, hi); }
MCT_HD MCT_FI int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return jmini(jmaxi(v, lo), hi); }

// Math.round(float) = (int) floor(a + 0.5) evaluated exactly in double.
MCT_HD MCT_FI int32_t jroundf(float a) { return d2i(floor(static_cast<double>(a) + 0.5)); }

// Mth.lerp(float t, a, b) = a + t * (b - a)
MCT_HD MCT_FI float lerpf(float t, float a, float b) {
    const float d = b - a;
    const float m = t * d;
    return a + m;
}

// Mth.lerp3 over the cube corners v[xyz]: x first, then y, then z.
MCT_HD MCT_FI float lerp3f(float tx, float ty, float tz, float v000, float v100, float v010, float v110, float v001, float v101,
                           float v011, float v111) {
    const float a = lerpf(ty, lerpf(tx, v000, v100), lerpf(tx, v010, v110));
    const float b = lerpf(ty, lerpf(tx, v001, v101), lerpf(tx, v011, v111));
    return lerpf(tz, a, b);
}

// Mth.smoothstep(t) = t^3 * (t * (t * 6 - 15) + 10), in float, in that association.
MCT_HD MCT_FI float smooth(float t) {
    const float t3 = (t * t) * t;
    float p = t * 6.0f;
    p = p - 15.0f;
    p = t * p;
    p = p + 10.0f;
    return t3 * p;
}

MCT_HD MCT_FI float cubef(float x) { return (x * x) * x; }

// ---- division by a constant ------------------------------------------------------------
// x / d from y = RN(1/d): the quotient q = RN(x * y), its remainder r = q * d - x (exact
// under one fused multiply-add) and one fused correction q - r * y. This is the one-FMA
// method of Brisebarre, Muller and Raina, "Accelerating correctly rounded floating-point
// division when the divisor is known in advance" (IEEE Trans. Computers 53(8), 2004). It is
// not correctly rounded for every x and d, so each caller names the input domain on which it
// equals the IEEE division bit for bit, checked exhaustively on the CPU
// (mcte_check_divisions). The remainder is taken as q * d - x, not x - q * d, so that a
// zero quotient keeps the sign of x. The device's IEEE division is about ten instructions,
// one of them at quarter rate; this is three.
MCT_HD MCT_FI float div_known(float x, float d, float y) {
    const float q = x * y;
    const float r = __builtin_fmaf(q, d, -x);
    return __builtin_fmaf(-r, y, q);
}
MCT_HD MCT_FI double div_known(double x, double d, double y) {
    const double q = x * y;
    const double r = __builtin_fma(q, d, -x);
    return __builtin_fma(-r, y, q);
}

// t / 24 as the squeeze takes it on the device: div_known where |t| >= 2^-120 and the IEEE
// division below, where the quotient is subnormal and div_known can be one unit off. Equal to
// t / 24.0f for every float t in [-1, 1], the squeeze's range. The division runs only in
// waves that hold such a t.
MCT_HD MCT_FI float squeeze_div24(float t) {
    const bool tiny = fabsf(t) < 0x1p-120f;
    float q = div_known(t, 24.0f, 1.0f / 24.0f);
#if defined(__HIP_DEVICE_COMPILE__)
    if (__builtin_amdgcn_ballot_w64(tiny)) {
        // an opaque copy, so that the division stays behind the branch instead of being
        // computed for every wave and selected
        float tt = t;
        asm volatile("" : "+v"(tt));
        if (tiny) q = tt / 24.0f;
    }
#else
    if (tiny) q = t / 24.0f;
#endif
    return q;
}

// ---- density-function unary ops (UnaryFunction) ----------------------------------
MCT_HD MCT_FI float op_squeeze(float x) {
    const float c = clampf(x, -1.0f, 1.0f);
#if defined(__HIP_DEVICE_COMPILE__)
    return c / 2.0f - squeeze_div24(cubef(c));
#else
    return c / 2.0f - cubef(c) / 24.0f;
#endif
}
MCT_HD MCT_FI float op_leaky(float x, float k) { return x > 0.0f ? x : x * k; }
// Mth.sqrt(float) = (float) Math.sqrt(x): the double sqrt of a float input, rounded to float.
MCT_HD MCT_FI float op_sqrt(float x) { return static_cast<float>(sqrt(static_cast<double>(x))); }
// Math.signum(float)
MCT_HD MCT_FI float op_sign(float x) { return (x != x || x == 0.0f) ? x : (x > 0.0f ? 1.0f : -1.0f); }

// Volume-path MIN/MAX: the buffered (first) value stays unless the second is strictly
// smaller / larger.
MCT_HD MCT_FI float vmin(float first, float second) { return second < first ? second : first; }
MCT_HD MCT_FI float vmax(float first, float second) { return second > first ? second : first; }

MCT_HD MCT_FI bool in_range(float v, float lo_incl, float hi_excl) { return v >= lo_incl && v < hi_excl; }

// CubicSpline.Multipoint between points (x1, y1) and (x2, y2) with derivatives d1, d2:
// t = (in - x1) / (x2 - x1); lerp(t, y1, y2) + t * (1 - t) * lerp(t, d1 * dx - dy, -d2 * dx + dy).
MCT_HD MCT_FI float spline_segment(float in, float x1, float x2, float d1, float d2, float y1, float y2) {
    const float t = (in - x1) / (x2 - x1);
    const float dx = x2 - x1, dy = y2 - y1;
    const float a = d1 * dx - dy;
    const float b = -d2 * dx + dy;
    const float l1 = lerpf(t, y1, y2);
    const float w = t * (1.0f - t);
    const float l2 = lerpf(t, a, b);
    return l1 + w * l2;
}

// Past an end point (x, y) with a non-zero derivative d: y + d * (in - x).
MCT_HD MCT_FI float spline_extend(float in, float x, float d, float y) {
    const float dx = in - x;
    return y + d * dx;
}

// GradientFunction (axis gradients with tiling). mode 0 clamp_to_edge, 1 repeat,
// 2 mirrored_repeat; factor = (toValue - fromValue) / (to - from) in float.
MCT_HD MCT_FI float gradient_at(int mode, int32_t from, int32_t lo, int32_t hi, int32_t range, float from_value, float factor, int32_t c) {
    if (mode == 0) return from_value + static_cast<float>(clampi(c, lo, hi) - from) * factor;
    const int32_t rel = iadd(c, -from);
    if (mode == 1) return from_value + static_cast<float>(fmod_(rel, range)) * factor;
    const int32_t tile = fdiv(rel, range);
    const int32_t local = rel - tile * range;
    return (tile & 1) == 0 ? from_value + static_cast<float>(local) * factor : from_value + static_cast<float>(range - local) * factor;
}

// InterpolatedFunction cell fill for one block from its 8 cell-corner values (v[xyz],
// x/z cell size 4 or 16, y cell size 8 or 1): lerp along z, then x, then y as the running
// sum of `step`, exactly as the volume path accumulates it.
// The cell's z then x lerps: its value at the bottom (lo) and the per-block y step.
MCT_HD MCT_FI void cell_lo_step(float v000, float v100, float v010, float v110, float v001, float v101, float v011, float v111, int xi, int zi,
                                float inv_xz, float inv_y, float& lo, float& step) {
    const float az = static_cast<float>(zi) * inv_xz;
    const float v00 = lerpf(az, v000, v001), v01 = lerpf(az, v010, v011);
    const float v10 = lerpf(az, v100, v101), v11 = lerpf(az, v110, v111);
    const float ax = static_cast<float>(xi) * inv_xz;
    lo = lerpf(ax, v00, v10);
    const float hi = lerpf(ax, v01, v11);
    step = (hi - lo) * inv_y;
}

// The value yi blocks above the bottom: lo + step * 0, then yi additions of the step.
MCT_HD MCT_FI float cell_at(float lo, float step, int yi) {
    float value = lo + step * 0.0f;
    for (int y = 0; y < yi; ++y) value += step;
    return value;
}

MCT_HD MCT_FI float cell_fill(float v000, float v100, float v010, float v110, float v001, float v101, float v011, float v111, int xi, int yi,
                              int zi, float inv_xz, float inv_y) {
    float lo, step;
    cell_lo_step(v000, v100, v010, v110, v001, v101, v011, v111, xi, zi, inv_xz, inv_y, lo, step);
    return cell_at(lo, step, yi);
}

}  // namespace mct
