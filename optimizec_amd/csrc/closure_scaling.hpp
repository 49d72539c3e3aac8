// closure_scaling.hpp -- the closure network's input / output scaling, shared by the
// training-pair gather and the coarse-run closure step.
//
// The network sees the small-scale part of the filtered fields: each channel minus its
// mean over the 5 x 5 box of grid points around the point (periodic), which removes the
// locally uniform advecting flow (Galilean invariance) and the large scales.  With
// w_s = rms of the small-scale vorticity and D = 2 pi / n, the inputs are
// (omega' / w_s, u' / (w_s D), w' / (w_s D)) and the target tau / w_s^2: the subgrid term
// of near-cutoff interactions scales as w_s^2, so a grid whose cutoff scales carry
// little vorticity gets a correspondingly small forcing.
#pragma once

#include <hip/hip_runtime.h>

namespace closure_scaling {

constexpr int kRadius = 2;

// f[y][x] - box mean, f an n x n periodic field
__device__ __forceinline__ float highpass(const float *__restrict__ f, int n, int y, int x)
{
    float s = 0.0f;
    for (int dy = -kRadius; dy <= kRadius; ++dy) {
        int yy = y + dy;
        yy += (yy < 0) ? n : ((yy >= n) ? -n : 0);
        const float *row = f + (size_t)yy * n;
        for (int dx = -kRadius; dx <= kRadius; ++dx) {
            int xx = x + dx;
            xx += (xx < 0) ? n : ((xx >= n) ? -n : 0);
            s += row[xx];
        }
    }
    constexpr float inv = 1.0f / float((2 * kRadius + 1) * (2 * kRadius + 1));
    return f[(size_t)y * n + x] - s * inv;
}

}  // namespace closure_scaling
