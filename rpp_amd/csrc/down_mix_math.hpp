// down_mix_math.hpp -- one down-mixed sample in the RPP HOST summation order
// (cpu/kernel/down_mixing.cpp:28-122), shared by the down-mix and the fused down-mix + resample
// kernels so both give the same bits.
#pragma once

namespace rppa {

// Channel weight 1/C, renormalised to unit sum when `normalize` (the HOST path's loop).
__device__ __forceinline__ float down_mix_weight(int C, int normalize)
{
    float w = 1.f / C;
    if (normalize) {
        float sum = 0.0f;
        for (int c = 0; c < C; c++) sum = sum + w;
        float inv = static_cast<float>(1.0 / sum);
        w = w * inv;
    }
    return w;
}

// sum_c x[c] * w: channels in groups of 4 (C < 8) or 8 (C >= 8) accumulated lane-wise, the lanes
// added as the HOST horizontal add does ((l0+l1)+(l2+l3); eight lanes
// ((l0+l4)+(l2+l6))+((l1+l5)+(l3+l7))), then the leftover channels one by one.
__device__ __forceinline__ float down_mix_sample(const float *x, int C, float w)
{
    if (C == 1) return x[0];
    int c = 0;
    float s;
    if (C > 7) {
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int aligned = (C / 8) * 8;
        for (; c < aligned; c += 8) {
#pragma unroll
            for (int l = 0; l < 8; l++) { float p = x[c + l] * w; acc[l] = acc[l] + p; }
        }
        float s0 = acc[0] + acc[4], s1 = acc[1] + acc[5], s2 = acc[2] + acc[6], s3 = acc[3] + acc[7];
        s = (s0 + s2) + (s1 + s3);
    } else {
        float acc[4] = {0, 0, 0, 0};
        const int aligned = (C / 4) * 4;
        for (; c < aligned; c += 4) {
#pragma unroll
            for (int l = 0; l < 4; l++) { float p = x[c + l] * w; acc[l] = acc[l] + p; }
        }
        s = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
    for (; c < C; c++) { float p = x[c] * w; s = s + p; }
    return s;
}

}  // namespace rppa
