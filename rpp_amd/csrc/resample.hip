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
// The table sits in LDS for the general path (its reads are data-dependent gathers).  For mono sources (and the fused
// down-mix, whose multi-channel input is reduced to mono first) with at most kMaxScale inputs per
// output, each wave first stages the input span its block reads into LDS.  When in/out is an
// integer k, every input position of a block is an integer, so each output whose taps are not
// clipped at the clip's ends reads its 2*lobes taps at the same window offsets -lobes .. lobes-1:
// those window values are looked up once per block and the output is a plain FIR in the HOST
// order (taps in fours, (s0+s2)+(s1+s3), then the leftover taps), bit for bit what the per-tap
// lookups give.
#include "common.hpp"
#include "down_mix_math.hpp"
#include "resample_math.hpp"

namespace rppa {
namespace {

constexpr int kThreads = 256;
constexpr int kBlockOut = 256;
constexpr int kMaxScale = 3;   // inputs per output up to which a block's input span is staged (sized so five
                               // blocks share a CU)
constexpr int kMaxTaps = 64;   // 2 * lobes up to which the span is staged
constexpr int kSpan = kBlockOut * kMaxScale + kMaxTaps + 8;   // staged floats per wave

// An integer-ratio block (every input position j*k exact): lane l takes outputs j = l + 64 i, so
// consecutive lanes read taps k floats apart (no LDS bank conflicts for odd k) and store
// consecutive outputs.  Clipped outputs at the clip's ends take the per-tap lookups.
template <int NT>
__device__ __forceinline__ void fir_block(const float *spB, const float *wts, const Win &win, float *y, long long ob,
                                          int outEnd, int k, int lobes, int inBlockRounded, int L, int lane)
{
    float wr[NT > 0 ? NT : 1];
    if (NT > 0) {
#pragma unroll
        for (int t = 0; t < NT; t++) wr[t] = wts[t];
    }
    for (int i = 0; i < kBlockOut / kWave; i++) {
        const int j = lane + kWave * i;
        const long long o = ob + j;
        if (o >= outEnd) break;
        const int xc = j * k;
        int loc0 = xc - lobes, loc1 = xc + lobes;
        float acc;
        if (loc0 + inBlockRounded >= 0 && loc1 + inBlockRounded <= L) {
            acc = NT > 0 ? fir_one<NT>(spB + loc0, wr, 0) : fir_one<0>(spB + loc0, wts, 2 * lobes);
        } else {
            if (loc0 + inBlockRounded < 0) loc0 = -inBlockRounded;
            if (loc1 + inBlockRounded > L) loc1 = L - inBlockRounded;
            acc = resample_one([&](int li) { return spB[li]; }, win, static_cast<float>(xc), loc0, loc1);
        }
        y[o] = acc;
    }
}

// DMIX: the source is interleaved multi-channel and is down-mixed (HOST order) before resampling;
// the output is mono.  Otherwise the drop-in resample: mono, or channels resampled independently.
template <bool DMIX>
__global__ void __launch_bounds__(kThreads)
resample_kernel(const float *__restrict__ src, uint32_t srcStride, float *__restrict__ dst, uint32_t dstStride,
                const float *__restrict__ inRates, const float *__restrict__ outRates, const int *__restrict__ dims,
                float wScale, float wCenter, int lobes, int lookupSize, const float *__restrict__ lutG, int normalize,
                float inRateAll = 0.0f, float outRateAll = 0.0f)
{
    extern __shared__ float lds[];
    float *lut = lds;
    float *wts = lds + (lookupSize + 3) / 4 * 4;          // FIR taps of an unclipped output
    float *span = wts + kMaxTaps;                          // [wave][kSpan] staged input
    const int b = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int C = max(dims[2 * b + 1], 1);
    const int L = min(dims[2 * b], static_cast<int>(srcStride / C));
    const int CO = DMIX ? 1 : C;                           // output channels
    const float ir = inRates ? inRates[b] : inRateAll, orr = outRates ? outRates[b] : outRateAll;
    const float *x = src + static_cast<size_t>(b) * srcStride;
    float *y = dst + static_cast<size_t>(b) * dstStride;
    const float dw = DMIX ? down_mix_weight(C, normalize) : 1.0f;
    auto sample = [&](int t) { return DMIX ? down_mix_sample(x + static_cast<size_t>(t) * C, C, dw) : x[t]; };
    const long long ob = (static_cast<long long>(blockIdx.x) * (kThreads / kWave) + wave) * kBlockOut;
    if (ir == orr) {  // no resampling: a copy of L*C samples (or of the L down-mixed samples)
        const long long n = min(static_cast<long long>(L), static_cast<long long>(dstStride / CO)) * CO;
        for (long long i = ob * CO + lane; i < min(n, (ob + kBlockOut) * CO); i += kWave)
            y[i] = DMIX ? sample(static_cast<int>(i)) : x[i];
        return;
    }
    const float lf = static_cast<float>(L) * orr;
    int outEnd = static_cast<int>(ceilf(lf / ir));
    outEnd = min(outEnd, static_cast<int>(dstStride / CO));
    const double scale = static_cast<double>(ir) / orr;
    const float fscale = static_cast<float>(scale);
    const int kInt = static_cast<int>(scale);
    // block-uniform choices (one clip per block)
    const bool staged = (DMIX || C == 1) && scale <= kMaxScale && 2 * lobes <= kMaxTaps;
    const bool fir = staged && static_cast<double>(kInt) == scale;
    // the table goes to LDS for the per-tap lookups of the general path; an integer-ratio block
    // looks up its 2*lobes window values once and reads the table in global memory only for the
    // outputs clipped at the clip's ends (the same values either way)
    if (!fir)
        for (int i = threadIdx.x; i < lookupSize; i += kThreads) lut[i] = lutG[i];
    const Win win{fir ? lutG : lut, wScale, wCenter, lookupSize};
    const int nGroups = (2 * lobes) / 4;
    if (fir && threadIdx.x < 2 * lobes) {
        const float p = static_cast<float>(static_cast<int>(threadIdx.x) - lobes);
        wts[threadIdx.x] = threadIdx.x < 4 * nGroups ? win.trunc_lookup(p) : win.floor_lookup(p);
    }
    const double inBlockRaw = static_cast<double>(ob) * scale;
    const int inBlockRounded = static_cast<int>(inBlockRaw);
    // staged span: inputs inBlockRounded + s0 .. + s0 + spanLen of this wave's block (zero outside
    // the clip, never read there); the float input positions drift from the exact ones by far
    // less than the two-sample margins
    const int s0 = -lobes - 2;
    const int spanLen = min(kSpan, static_cast<int>(ceil(kBlockOut * scale)) + 2 * lobes + 4);
    float *sp = span + wave * kSpan;
    if (staged && ob < outEnd) {
        // sixteen samples per lane in flight: the loads of a batch are issued before any is used
        // (mono and stereo; other channel counts down-mix sample by sample)
        constexpr int kU = 16;
        const bool pairs = C == 2 && (reinterpret_cast<uintptr_t>(x) & 7) == 0;   // 8-byte stereo loads
        if (C == 1 || pairs) {
            for (int j0 = lane; j0 < spanLen; j0 += kU * kWave) {
                float v[kU][2];
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    const int t = min(max(inBlockRounded + s0 + j0 + u * kWave, 0), L - 1);
                    if (pairs) {
                        const float2 p2 = *reinterpret_cast<const float2 *>(x + 2 * static_cast<size_t>(t));
                        v[u][0] = p2.x;
                        v[u][1] = p2.y;
                    } else {
                        v[u][0] = x[t];
                        v[u][1] = 0.0f;
                    }
                }
#pragma unroll
                for (int u = 0; u < kU; u++) {
                    const int j = j0 + u * kWave, t = inBlockRounded + s0 + j;
                    if (j < spanLen) sp[j] = (t >= 0 && t < L) ? (pairs ? down_mix_sample(v[u], 2, dw) : v[u][0]) : 0.0f;
                }
            }
        } else {
            for (int j = lane; j < spanLen; j += kWave) {
                const int t = inBlockRounded + s0 + j;
                sp[j] = (t >= 0 && t < L) ? sample(t) : 0.0f;
            }
        }
    }
    __syncthreads();                                       // table, taps and spans staged
    if (ob >= outEnd) return;
    if (fir) {
        const float *spB = span + wave * kSpan - s0;
        if (lobes == 16) fir_block<32>(spB, wts, win, y, ob, outEnd, kInt, lobes, inBlockRounded, L, lane);
        else fir_block<0>(spB, wts, win, y, ob, outEnd, kInt, lobes, inBlockRounded, L, lane);
        return;
    }
    float inPos = static_cast<float>(inBlockRaw - inBlockRounded);
    for (int j = 0; j < 4 * lane; j++) inPos += fscale;
    const float *inBlock = x + static_cast<size_t>(inBlockRounded) * C;
    const float *spB = sp - s0;                            // spB[li] = input inBlockRounded + li
    for (int i = 0; i < 4; i++, inPos += fscale) {
        const long long o = ob + 4 * lane + i;
        if (o >= outEnd) break;
        const int xc = static_cast<int>(ceilf(inPos));
        int loc0 = xc - lobes, loc1 = xc + lobes;
        if (loc0 + inBlockRounded < 0) loc0 = -inBlockRounded;
        if (loc1 + inBlockRounded > L) loc1 = L - inBlockRounded;
        if (staged) {
            y[o] = resample_one([&](int li) { return spB[li]; }, win, inPos, loc0, loc1);
        } else if (DMIX || C == 1) {
            y[o] = resample_one([&](int li) { return sample(inBlockRounded + li); }, win, inPos, loc0, loc1);
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

size_t resample_lds_bytes(int lookupSize)
{
    return sizeof(float) * ((lookupSize + 3) / 4 * 4 + kMaxTaps + (kThreads / kWave) * kSpan);
}

// Each clip's resampled length, ceil(samples * outRate / inRate) in FP32 as the resample kernel
// computes its output count.
__global__ void resampled_length_kernel(const int *__restrict__ dims, uint32_t srcStride, float inRate, float outRate,
                                        int B, int *__restrict__ lengths)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int C = max(dims[2 * b + 1], 1);
    const int L = max(min(dims[2 * b], static_cast<int>(srcStride / C)), 0);
    const float lf = static_cast<float>(L) * outRate;
    lengths[b] = static_cast<int>(ceilf(lf / inRate));
}

}  // namespace

hipError_t down_mix_resample_launch(hipStream_t st, const float *src, uint32_t srcStride, float *dst,
                                    uint32_t dstStride, unsigned B, const int *dims, float inRate, float outRate,
                                    const RpptResamplingWindow &w, int normalize, int *lengths)
{
    dim3 grid(ceil_div(dstStride, kBlockOut * (kThreads / kWave)), B);
    const size_t bytes = resample_lds_bytes(w.lookupSize);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&resample_kernel<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
    hipLaunchKernelGGL(resample_kernel<true>, grid, dim3(kThreads), bytes, st, src, srcStride, dst, dstStride, nullptr,
                       nullptr, dims, w.scale, w.center, w.lobes, w.lookupSize, w.lookupPinned, normalize, inRate,
                       outRate);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(resampled_length_kernel, dim3(ceil_div(B, 256u)), dim3(256), 0, st, dims, srcStride, inRate,
                       outRate, static_cast<int>(B), lengths);
    return hipGetLastError();
}

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
    const size_t bytes = rppa::resample_lds_bytes(window.lookupSize);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&rppa::resample_kernel<false>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
    hipLaunchKernelGGL(rppa::resample_kernel<false>, grid, dim3(rppa::kThreads), bytes, rppHandle->stream,
                       static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, static_cast<float *>(dstPtr),
                       dstDescPtr->strides.nStride, inRateTensor, outRateTensor, srcDimsTensor, window.scale,
                       window.center, window.lobes, window.lookupSize, window.lookupPinned, 0);
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}

// Down-mix (rppt_down_mixing_gpu's arithmetic) and resample (rppt_resample_gpu's) in one launch:
// src interleaved [n][samples * channels], srcDimsTensor[2b] = samples, [2b+1] = channels; dst
// mono [n][nStride].  Bit-identical to the two calls in sequence; the down-mixed signal never
// goes to memory.
RPPA_EXPORT RppStatus rppt_down_mix_resample_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                                 RpptDescPtr dstDescPtr, Rpp32f *inRateTensor, Rpp32f *outRateTensor,
                                                 Rpp32s *srcDimsTensor, bool normalizeWeights,
                                                 RpptResamplingWindow &window, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    const size_t dims = srcDescPtr->numDims - 1;
    if (dims != 1 && dims != 2) return RPP_ERROR_INVALID_SRC_DIMS;
    if (srcDescPtr->dataType != F32 || dstDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !inRateTensor || !outRateTensor || !srcDimsTensor || !window.lookupPinned)
        return RPP_ERROR_INVALID_ARGUMENTS;
    if (window.lookupSize < 2 || window.lookupSize > 16384 || window.lobes < 0) return RPP_ERROR_INVALID_ARGUMENTS;
    const unsigned B = srcDescPtr->n;
    if (B == 0) return RPP_SUCCESS;
    const unsigned outFrames = dstDescPtr->strides.nStride;
    dim3 grid(rppa::ceil_div(outFrames, rppa::kBlockOut * (rppa::kThreads / rppa::kWave)), B);
    const size_t bytes = rppa::resample_lds_bytes(window.lookupSize);
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&rppa::resample_kernel<true>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
    hipLaunchKernelGGL(rppa::resample_kernel<true>, grid, dim3(rppa::kThreads), bytes, rppHandle->stream,
                       static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, static_cast<float *>(dstPtr),
                       dstDescPtr->strides.nStride, inRateTensor, outRateTensor, srcDimsTensor, window.scale,
                       window.center, window.lobes, window.lookupSize, window.lookupPinned, normalizeWeights ? 1 : 0);
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}
