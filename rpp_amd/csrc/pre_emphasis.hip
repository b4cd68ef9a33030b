// pre_emphasis.hip -- rppt_pre_emphasis_filter_gpu (api/rppt_tensor_audio_augmentations.h:156).
//
// y[0] = x[0] - c*border (border: CLAMP x[0], REFLECT x[1], ZERO 0), y[t] = x[t] - c*x[t-1]
// for t < srcLengthTensor[b] (cpu/kernel/pre_emphasis_filter.cpp:27-75).  The product is
// rounded before the subtraction (no FMA), matching the reference's -mno-fma build, so the
// output is bit-identical to RPP HOST.  HBM-bound: 8 bytes per sample.
#include "common.hpp"

namespace rppa {
namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 4;

__global__ void __launch_bounds__(kThreads)
pre_emphasis_kernel(const float *__restrict__ src, uint32_t srcStride, float *__restrict__ dst, uint32_t dstStride,
                    const int *__restrict__ lengths, const float *__restrict__ coeffs, int border)
{
    const int b = blockIdx.y;
    // a length beyond either stride would run into the next clip; clamp it to the buffer
    const int L = min(lengths[b], static_cast<int>(min(srcStride, dstStride)));
    const size_t t0 = (static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x) * kPerThread;
    if (t0 >= static_cast<size_t>(L)) return;
    const float *x = src + static_cast<size_t>(b) * srcStride;
    float *y = dst + static_cast<size_t>(b) * dstStride;
    const float c = coeffs[b];
    float prev;
    if (t0 == 0) prev = (border == CLAMP) ? x[0] : (border == REFLECT && srcStride > 1 ? x[1] : 0.0f);
    else prev = x[t0 - 1];
#pragma unroll
    for (int i = 0; i < kPerThread; i++) {
        size_t t = t0 + i;
        if (t >= static_cast<size_t>(L)) break;
        float cur = x[t];
        float p = c * prev;
        y[t] = cur - p;
        prev = cur;
    }
}

}  // namespace
}  // namespace rppa

// Validation follows rppt_tensor_audio_augmentations.cpp:452-484.
RPPA_EXPORT RppStatus rppt_pre_emphasis_filter_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                                   RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor,
                                                   Rpp32f *coeffTensor, RpptAudioBorderType borderType,
                                                   rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (srcDescPtr->numDims != 2) return RPP_ERROR_INVALID_SRC_DIMS;
    if (srcDescPtr->dataType != F32 || dstDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcLengthTensor || !coeffTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    const unsigned B = srcDescPtr->n;
    if (B == 0) return RPP_SUCCESS;
    const uint32_t maxLen = srcDescPtr->strides.nStride;
    dim3 grid(rppa::ceil_div(maxLen, rppa::kThreads * rppa::kPerThread), B);
    hipLaunchKernelGGL(rppa::pre_emphasis_kernel, grid, dim3(rppa::kThreads), 0, rppHandle->stream,
                       static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, static_cast<float *>(dstPtr),
                       dstDescPtr->strides.nStride, srcLengthTensor, coeffTensor, static_cast<int>(borderType));
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}
