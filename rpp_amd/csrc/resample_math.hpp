// resample_math.hpp -- the windowed-sinc arithmetic of one resampled output in the RPP HOST order
// (cpu/kernel/resample.cpp:27-146), shared by the resample kernels and the fused front end of the
// log-mel launch so every path gives the same bits.
#pragma once

namespace rppa {

struct Win {
    const float *lut;
    float scale, center;
    int lookupSize;
    __device__ float floor_lookup(float x) const
    {
        float locRaw = x * scale + center;
        int i = static_cast<int>(floorf(locRaw));
        float weight = locRaw - static_cast<float>(i);
        i = max(min(i, lookupSize - 2), 0);
        float cur = lut[i], nxt = lut[i + 1];
        return cur + weight * (nxt - cur);
    }
    __device__ float trunc_lookup(float x) const
    {
        float locRaw = x * scale + center;
        int i = static_cast<int>(locRaw);
        float weight = locRaw - static_cast<float>(i);
        i = max(min(i, lookupSize - 2), 0);  // in range for every tap the HOST path visits
        float cur = lut[i], nxt = lut[i + 1];
        return cur + weight * (nxt - cur);
    }
};

// One mono output in the HOST order: taps [loc0, loc1) of the block's input `in` (indexable by
// the tap index relative to the block start), window at li - inPos.
template <typename In>
__device__ __forceinline__ float resample_one(const In &in, const Win &win, float inPos, int loc0, int loc1)
{
    int li = loc0;
    const float locBegin = static_cast<float>(li) - inPos;
    float p0 = locBegin + 0.f, p1 = locBegin + 1.f, p2 = locBegin + 2.f, p3 = locBegin + 3.f;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    for (; li + 3 < loc1; li += 4) {
        float q;
        q = in(li) * win.trunc_lookup(p0);     a0 = a0 + q;
        q = in(li + 1) * win.trunc_lookup(p1); a1 = a1 + q;
        q = in(li + 2) * win.trunc_lookup(p2); a2 = a2 + q;
        q = in(li + 3) * win.trunc_lookup(p3); a3 = a3 + q;
        p0 = p0 + 4.0f; p1 = p1 + 4.0f; p2 = p2 + 4.0f; p3 = p3 + 4.0f;
    }
    float acc = (a0 + a2) + (a1 + a3);
    float xx = static_cast<float>(li) - inPos;
    for (; li < loc1; li++, xx += 1.0f) {
        float q = in(li) * win.floor_lookup(xx);
        acc = acc + q;
    }
    return acc;
}

// Unclipped output of an integer-ratio block: taps xt[0 .. nt) against the block's window values
// w[0 .. nt) in the HOST order (fours, (s0+s2)+(s1+s3), then the leftover taps).  NT > 0 fixes
// the tap count at compile time so the window values sit in registers (W is then a register
// array), NT == 0 reads them from LDS.
template <int NT, typename Wv>
__device__ __forceinline__ float fir_one(const float *xt, const Wv &w, int nt)
{
    const int n = NT > 0 ? NT : nt;
    const int ng = n / 4;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
#pragma unroll
    for (int g = 0; g < (NT > 0 ? NT / 4 : ng); g++) {
        float q;
        q = xt[4 * g] * w[4 * g];         a0 = a0 + q;
        q = xt[4 * g + 1] * w[4 * g + 1]; a1 = a1 + q;
        q = xt[4 * g + 2] * w[4 * g + 2]; a2 = a2 + q;
        q = xt[4 * g + 3] * w[4 * g + 3]; a3 = a3 + q;
    }
    float acc = (a0 + a2) + (a1 + a3);
#pragma unroll
    for (int t = 4 * ng; t < n; t++) { float q = xt[t] * w[t]; acc = acc + q; }
    return acc;
}

}  // namespace rppa
