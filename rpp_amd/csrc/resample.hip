// resample.hip -- rppt_resample_gpu (api/rppt_tensor_audio_augmentations.h:310).
//
// Windowed-sinc resampling with the caller's lookup table (RpptResamplingWindow,
// api/rppdefs.h:751-790), following cpu/kernel/resample.cpp:27-146: outEnd = ceil(L*out/in) in
// FP32; outputs in blocks of 256 whose input position starts from the FP64 block base and then
// advances by the FP32 step in->out once per output; taps [ceil(x) - lobes, ceil(x) + lobes)
// clipped to the clip; the window value is the linear interpolation of the table.  For mono
// input the taps go in groups of four with the truncating table lookup, the four partial sums
// are combined as (s0+s2)+(s1+s3), and the leftover taps use the flooring, clamped lookup -- the
// HOST path's exact order, so results are bit-identical to RPP HOST.
//
// Work split: one wave per 256-output block; lane l owns outputs 4l..4l+3 of the block and
// reaches its first input position by the same sequence of FP32 additions the HOST loop makes.
// The table sits in LDS (its reads are data-dependent gathers).
#include "common.hpp"

namespace rppa {
namespace {

constexpr int kThreads = 256;
constexpr int kBlockOut = 256;

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

__global__ void __launch_bounds__(kThreads)
resample_kernel(const float *__restrict__ src, uint32_t srcStride, float *__restrict__ dst, uint32_t dstStride,
                const float *__restrict__ inRates, const float *__restrict__ outRates, const int *__restrict__ dims,
                float wScale, float wCenter, int lobes, int lookupSize, const float *__restrict__ lutG)
{
    extern __shared__ float lut[];
    for (int i = threadIdx.x; i < lookupSize; i += kThreads) lut[i] = lutG[i];
    __syncthreads();
    const Win win{lut, wScale, wCenter, lookupSize};
    const int b = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int C = max(dims[2 * b + 1], 1);
    const int L = min(dims[2 * b], static_cast<int>(srcStride / C));
    const float ir = inRates[b], orr = outRates[b];
    const float *x = src + static_cast<size_t>(b) * srcStride;
    float *y = dst + static_cast<size_t>(b) * dstStride;
    const long long ob = (static_cast<long long>(blockIdx.x) * (kThreads / kWave) + wave) * kBlockOut;
    if (ir == orr) {  // plain copy of L*C samples
        const long long n = min(static_cast<long long>(L), static_cast<long long>(dstStride / C)) * C;
        for (long long i = ob * C + lane; i < min(n, (ob + kBlockOut) * C); i += kWave) y[i] = x[i];
        return;
    }
    const float lf = static_cast<float>(L) * orr;
    int outEnd = static_cast<int>(ceilf(lf / ir));
    outEnd = min(outEnd, static_cast<int>(dstStride / C));
    if (ob >= outEnd) return;
    const double scale = static_cast<double>(ir) / orr;
    const float fscale = static_cast<float>(scale);
    const double inBlockRaw = static_cast<double>(ob) * scale;
    const int inBlockRounded = static_cast<int>(inBlockRaw);
    float inPos = static_cast<float>(inBlockRaw - inBlockRounded);
    for (int j = 0; j < 4 * lane; j++) inPos += fscale;
    const float *inBlock = x + static_cast<size_t>(inBlockRounded) * C;
    for (int i = 0; i < 4; i++, inPos += fscale) {
        const long long o = ob + 4 * lane + i;
        if (o >= outEnd) break;
        const int xc = static_cast<int>(ceilf(inPos));
        int loc0 = xc - lobes, loc1 = xc + lobes;
        if (loc0 + inBlockRounded < 0) loc0 = -inBlockRounded;
        if (loc1 + inBlockRounded > L) loc1 = L - inBlockRounded;
        if (C == 1) {
            int li = loc0;
            const float locBegin = static_cast<float>(li) - inPos;
            float p0 = locBegin + 0.f, p1 = locBegin + 1.f, p2 = locBegin + 2.f, p3 = locBegin + 3.f;
            float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
            for (; li + 3 < loc1; li += 4) {
                float q;
                q = inBlock[li] * win.trunc_lookup(p0);     a0 = a0 + q;
                q = inBlock[li + 1] * win.trunc_lookup(p1); a1 = a1 + q;
                q = inBlock[li + 2] * win.trunc_lookup(p2); a2 = a2 + q;
                q = inBlock[li + 3] * win.trunc_lookup(p3); a3 = a3 + q;
                p0 = p0 + 4.0f; p1 = p1 + 4.0f; p2 = p2 + 4.0f; p3 = p3 + 4.0f;
            }
            float acc = (a0 + a2) + (a1 + a3);
            float xx = static_cast<float>(li) - inPos;
            for (; li < loc1; li++, xx += 1.0f) {
                float q = inBlock[li] * win.floor_lookup(xx);
                acc = acc + q;
            }
            y[o] = acc;
        } else {
            float acc[16];
            const int Cc = min(C, 16);
            for (int c = 0; c < 16; c++) acc[c] = 0.0f;
            float xx = static_cast<float>(loc0) - inPos;
            for (int li = loc0; li < loc1; li++, xx += 1.0f) {
                const float w = win.floor_lookup(xx);
#pragma unroll
                for (int c = 0; c < 16; c++)
                    if (c < Cc) { float q = inBlock[static_cast<size_t>(li) * C + c] * w; acc[c] = acc[c] + q; }
            }
#pragma unroll
            for (int c = 0; c < 16; c++)
                if (c < Cc) y[o * C + c] = acc[c];
        }
    }
}

}  // namespace
}  // namespace rppa

// Validation follows rppt_tensor_audio_augmentations.cpp:539-575.
RPPA_EXPORT RppStatus rppt_resample_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                        RpptDescPtr dstDescPtr, Rpp32f *inRateTensor, Rpp32f *outRateTensor,
                                        Rpp32s *srcDimsTensor, RpptResamplingWindow &window, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    const size_t dims = srcDescPtr->numDims - 1;
    if (dims != 1 && dims != 2) return RPP_ERROR_INVALID_SRC_DIMS;
    if (srcDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !inRateTensor || !outRateTensor || !srcDimsTensor || !window.lookupPinned)
        return RPP_ERROR_INVALID_ARGUMENTS;
    if (window.lookupSize < 2 || window.lookupSize > 16384 || window.lobes < 0) return RPP_ERROR_INVALID_ARGUMENTS;
    const unsigned B = srcDescPtr->n;
    if (B == 0) return RPP_SUCCESS;
    const unsigned outFrames = dstDescPtr->strides.nStride;  // upper bound on outputs per clip
    dim3 grid(rppa::ceil_div(outFrames, rppa::kBlockOut * (rppa::kThreads / rppa::kWave)), B);
    hipLaunchKernelGGL(rppa::resample_kernel, grid, dim3(rppa::kThreads), sizeof(float) * window.lookupSize,
                       rppHandle->stream, static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride,
                       static_cast<float *>(dstPtr), dstDescPtr->strides.nStride, inRateTensor, outRateTensor,
                       srcDimsTensor, window.scale, window.center, window.lobes, window.lookupSize,
                       window.lookupPinned);
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}
