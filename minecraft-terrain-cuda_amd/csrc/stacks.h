// Noise stacks: the octave loops of NormalNoise (two PerlinNoise stacks per kept octave, the
// second at frequency x 1.0181268882175227) and of BlendedNoise's fbm sets, as templates over
// the layer range [First, First + Count) of the layer tables. The generated evaluators call
// one instantiation per noise key (or blended set) and path:
//
//   stack_volume(_smeared)      volume path at block (bx, by, bz): per layer
//                               acc = acc + (amplitude * amp) * sample(block * (scale * freq))
//   stack_volume_tab(_smeared)  the same with the x / z axis records of each layer read from
//                               the corner stage's per-column table
//   stack_volume_tab_y          the same with the y records read from the seed's corner-row
//                               table too
//   stack_point(_smeared)       point path at (x, y, z) already scaled: value += amp * sample
//
// (reference: cpu/include/mccuda/noise_eval.h:267-304, cpu/src/seed_tables.cpp:106-184).
// The loops are kept out of line and not unrolled: inlined into the evaluators, their octave
// temporaries pile onto the evaluators' values and force register spills.
//
// On the device a stack runs once per distinct octave table among the wave's lanes: the
// lanes of the first active lane's table run it, reading permutations through a buffer
// resource of that table (one pass when the wave shares a seed), then the rest loop again.
#pragma once

#include "noise.h"
#include "gen/model.h"
#include "gen/opcounts.h"

#define MCT_OPS_OF(k) MCT_OPS((k).f64, (k).f32, (k).i32, (k).mem)

namespace mct {

template <int First, int Count, class Sample>
MCT_HD MCT_FI float octave_sum(const Octave* O, Sample sample) {
    float acc = 0.0f;
#if defined(__HIP_DEVICE_COMPILE__)
    for (;;) {
        const Octave* T = wave_first(O);
        if (T == O) {
            const __amdgpu_buffer_rsrc_t R = octave_rsrc(T, mctgen::kNumLayers);
#pragma clang loop unroll(disable)
            for (int i = First; i < First + Count; ++i) acc = sample(acc, T[i], PermBuf(R, i), i);
            break;
        }
    }
#else
    for (int i = First; i < First + Count; ++i) acc = sample(acc, O[i], PermRef{O[i].q}, i);
#endif
    return acc;
}

template <int First, int Count>
MCT_HD MCT_NI float stack_volume(const Octave* O, int32_t bx, int32_t by, int32_t bz, double xz, double ys, float amplitude) {
    return octave_sum<First, Count>(O, [&](float acc, const Octave& o, const auto& P, int i) MCT_CLOSURE {
        MCT_OPS_OF(mctgen::kOctaveVolume);
        return acc + (amplitude * mctgen::kLayerAmp[i]) * perlin_volume(o, P, bx, by, bz, xz * mctgen::kLayerFreq[i], ys * mctgen::kLayerFreq[i]);
    });
}

template <int First, int Count>
MCT_HD MCT_NI float stack_volume_smeared(const Octave* O, int32_t bx, int32_t by, int32_t bz, double xz, double ys, float amplitude) {
    return octave_sum<First, Count>(O, [&](float acc, const Octave& o, const auto& P, int i) MCT_CLOSURE {
        MCT_OPS_OF(mctgen::kOctaveVolume);
        MCT_OPS_OF(mctgen::kSmearExtra);
        return acc + (amplitude * mctgen::kLayerAmp[i]) *
                         smeared_volume(o, P, mctgen::kLayerSmear[i], bx, by, bz, xz * mctgen::kLayerFreq[i], ys * mctgen::kLayerFreq[i]);
    });
}

// A: the table records of the stack's layers, in layer order (A[i - First] for layer i).
template <int First, int Count>
MCT_HD MCT_NI float stack_volume_tab(const Octave* O, const MCT_LDS AxisXZ* A, int32_t by, double ys, float amplitude) {
    return octave_sum<First, Count>(O, [&](float acc, const Octave& o, const auto& P, int i) MCT_CLOSURE {
        MCT_OPS_OF(mctgen::kOctaveVolumeTab);
        const AxisXZ a = A[i - First];
        const Axis Y = axis_plain(by, ys * mctgen::kLayerFreq[i], o.oy);
        return acc + (amplitude * mctgen::kLayerAmp[i]) * volume_sample_xz(P, a, Y);
    });
}

template <int First, int Count>
MCT_HD MCT_NI float stack_volume_tab_smeared(const Octave* O, const MCT_LDS AxisXZ* A, int32_t by, double ys, float amplitude) {
    return octave_sum<First, Count>(O, [&](float acc, const Octave& o, const auto& P, int i) MCT_CLOSURE {
        MCT_OPS_OF(mctgen::kOctaveVolumeTab);
        MCT_OPS_OF(mctgen::kSmearExtra);
        const AxisXZ a = A[i - First];
        const Axis Y = axis_smeared(by, ys * mctgen::kLayerFreq[i], o.oy, mctgen::kLayerSmear[i]);
        return acc + (amplitude * mctgen::kLayerAmp[i]) * volume_sample_xz(P, a, Y);
    });
}

// The same with the y axis records read too, from the seed's corner-row table: Yt[(i - First)
// * Rows] is layer i's record at the corner's row (the records stack_volume_tab(_smeared)
// computes, tabulated per seed by corner_y_axis). On the device the gradients come from the
// LDS table (the corner kernel fills it).
template <int First, int Count, int Rows>
MCT_HD MCT_NI float stack_volume_tab_y(const Octave* O, const MCT_LDS AxisXZ* A, const Axis* Yt, float amplitude) {
    return octave_sum<First, Count>(O, [&](float acc, const Octave& o, const auto& P, int i) MCT_CLOSURE {
        MCT_OPS_OF(mctgen::kOctaveVolumeTabY);
        (void)o;
        const AxisXZ a = A[i - First];
        const Axis Y = Yt[(i - First) * Rows];
        return acc + (amplitude * mctgen::kLayerAmp[i]) * volume_sample_xz_tab(P, a, Y);
    });
}

template <int First, int Count>
MCT_HD MCT_NI float stack_point(const Octave* O, double x, double y, double z) {
    return octave_sum<First, Count>(O, [&](float value, const Octave& o, const auto& P, int i) MCT_CLOSURE {
        MCT_OPS_OF(mctgen::kOctavePoint);
        const double f = mctgen::kLayerFreq[i];
        return value + mctgen::kLayerAmp[i] * perlin_point(o, P, x * f, y * f, z * f);
    });
}

template <int First, int Count>
MCT_HD MCT_NI float stack_point_smeared(const Octave* O, double x, double y, double z) {
    return octave_sum<First, Count>(O, [&](float value, const Octave& o, const auto& P, int i) MCT_CLOSURE {
        MCT_OPS_OF(mctgen::kOctavePoint);
        MCT_OPS_OF(mctgen::kSmearExtra);
        const double f = mctgen::kLayerFreq[i];
        return value + mctgen::kLayerAmp[i] * smeared_point(o, P, mctgen::kLayerSmear[i], x * f, y * f, z * f);
    });
}

// One record of the corner stage's column table: the x and z axes of one layer at block
// column (x, z) with xz factor xs = scale * frequency, the x record carrying the first-level
// permutation entries p[x & 255] | p[(x & 255) + 1] << 16 in place of its cell.
MCT_HD MCT_FI AxisXZ site_record(const Octave& o, int32_t x, int32_t z, double xs) {
    MCT_OPS_OF(mctgen::kSiteRecord);
    Axis X = axis_plain(x, xs, o.ox);
    const int xi = X.f & 0xFF;
    X.f = static_cast<int32_t>(o.q[xi] & 0xFF) | (static_cast<int32_t>(o.q[xi] >> 8) << 16);
    return AxisXZ{X, axis_plain(z, xs, o.oz)};
}

}  // namespace mct
