// mel_filter_bank.hip -- rppt_mel_filter_bank_gpu (api/rppt_tensor_audio_augmentations.h:274), FP32.
//
// out[m][f] = sum_k W[m][k] * S[k][f] over the triangular filters of
// cpu/kernel/mel_filter_bank.cpp:27-198.  The tables come from tables.hip (built on the GPU with
// the reference's algorithm).  One thread owns one (filter, frame) output and walks the
// filter's bins in ascending order, adding the rounded product each time (no FMA) starting
// from 0 -- the reference's order of operations, so given the same spectrogram and tables the
// result is bit-identical to RPP HOST.  Consecutive threads take consecutive frames, so both
// the NFT spectrogram reads and the NFT mel writes are coalesced.
#include "common.hpp"

namespace rppa {
namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads)
mel_gather_kernel(const float *__restrict__ src, uint32_t srcNStride, uint32_t srcHStride,
                  float *__restrict__ dst, uint32_t dstNStride, uint32_t dstHStride, const int *__restrict__ srcDims,
                  const float *__restrict__ weightsDown, const int *__restrict__ intervals,
                  const float *__restrict__ normFactors, const int *__restrict__ filterBins, int normalize)
{
    const int b = blockIdx.z;
    const int m = blockIdx.y;
    const int f = blockIdx.x * kThreads + threadIdx.x;
    const int F = min(srcDims[2 * b + 1], static_cast<int>(min(srcHStride, dstHStride)));
    if (f >= F) return;
    const float *s = src + static_cast<size_t>(b) * srcNStride + f;
    const float norm = normFactors[m];
    const int k0 = filterBins[2 * m], k1 = filterBins[2 * m + 1];
    float acc = 0.0f;
    for (int k = k0; k < k1; k++) {
        const float wd = weightsDown[k];
        float w = (intervals[k] == m) ? 1.0f - wd : wd;   // rising slope of m, else falling slope
        if (normalize) w = w * norm;
        const float p = s[static_cast<size_t>(k) * srcHStride] * w;
        acc = acc + p;
    }
    dst[static_cast<size_t>(b) * dstNStride + static_cast<size_t>(m) * dstHStride + f] = acc;
}

}  // namespace
}  // namespace rppa

// Validation follows rppt_tensor_audio_augmentations.cpp:488-535.  numBins is srcDesc->h, as the
// reference HIP path takes it (hip/kernel/mel_filter_bank.cpp:163-167).
RPPA_EXPORT RppStatus rppt_mel_filter_bank_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                               RpptDescPtr dstDescPtr, Rpp32s *srcDims, Rpp32f maxFreq, Rpp32f minFreq,
                                               RpptMelScaleFormula melFormula, Rpp32s numFilter, Rpp32f sampleRate,
                                               bool normalize, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (srcDescPtr->numDims - 1 != 2) return RPP_ERROR_INVALID_SRC_DIMS;
    if (srcDescPtr->layout != NFT) return RPP_ERROR_INVALID_SRC_LAYOUT;
    if (dstDescPtr->layout != NFT) return RPP_ERROR_INVALID_DST_LAYOUT;
    if (maxFreq < 0 || maxFreq > sampleRate / 2) return RPP_ERROR_INVALID_ARGUMENTS;
    if (minFreq < 0 || minFreq > sampleRate / 2) return RPP_ERROR_INVALID_ARGUMENTS;
    if (srcDescPtr->dataType != F32 || dstDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcDims || numFilter <= 0 || srcDescPtr->h < 2) return RPP_ERROR_INVALID_ARGUMENTS;
    const unsigned B = srcDescPtr->n;
    const unsigned maxFrames = dstDescPtr->w;
    if (B == 0 || maxFrames == 0) return RPP_SUCCESS;
    const rppa::MelTables *t = rppa::get_mel_tables(rppHandle, static_cast<int>(srcDescPtr->h), sampleRate, maxFreq,
                                                    minFreq, static_cast<int>(melFormula), numFilter, normalize ? 1 : 0);
    if (!t) return RPP_ERROR;
    dim3 grid(rppa::ceil_div(maxFrames, rppa::kThreads), numFilter, B);
    hipLaunchKernelGGL(rppa::mel_gather_kernel, grid, dim3(rppa::kThreads), 0, rppHandle->stream,
                       static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, srcDescPtr->strides.hStride,
                       static_cast<float *>(dstPtr), dstDescPtr->strides.nStride, dstDescPtr->strides.hStride, srcDims,
                       t->weightsDown, t->intervals, t->normFactors, t->filterBins, normalize ? 1 : 0);
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}
