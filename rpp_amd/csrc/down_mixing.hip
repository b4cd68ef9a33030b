// down_mixing.hip -- rppt_down_mixing_gpu (api/rppt_tensor_audio_augmentations.h:188).
//
// mono[t] = sum_c x[t*C + c] * w_c with w_c = 1/C (renormalised to unit sum when
// normalizeWeights), srcDimsTensor[2b] = samples, [2b+1] = channels, interleaved input
// (cpu/kernel/down_mixing.cpp:28-122).  The summation order is the HOST path's: channels in
// groups of 4 (C < 8) or 8 (C >= 8) accumulate lane-wise, the lanes are added as its horizontal
// add does ((l0+l1)+(l2+l3); eight lanes ((l0+l4)+(l2+l6))+((l1+l5)+(l3+l7))), then the leftover
// channels one by one -- so the output is bit-identical to RPP HOST.  (RPP's HIP kernel ignores
// normalizeWeights and sums in a different order, hip/kernel/down_mixing.cpp:28-70.)
// One thread per output sample; HBM-bound: 4*(C+1) bytes per sample.
#include "common.hpp"
#include "down_mix_math.hpp"

namespace rppa {
namespace {

constexpr int kThreads = 256;

// Stereo rows whose source and destination are 16-byte aligned: four outputs per thread from two
// 16-byte loads and one 16-byte store (the same arithmetic per sample).
__global__ void __launch_bounds__(kThreads)
down_mix_stereo4_kernel(const float *__restrict__ src, uint32_t srcStride, float *__restrict__ dst,
                        uint32_t dstStride, const int *__restrict__ dims, int normalize)
{
    const int b = blockIdx.y;
    if (dims[2 * b + 1] != 2) return;   // other rows: down_mix_kernel
    const long long S = min(static_cast<long long>(dims[2 * b]),
                            min(static_cast<long long>(srcStride) / 2, static_cast<long long>(dstStride)));
    const long long t = (static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x) * 4;
    if (t >= S) return;
    const float *x = src + static_cast<size_t>(b) * srcStride + 2 * t;
    float *y = dst + static_cast<size_t>(b) * dstStride + t;
    const float w = down_mix_weight(2, normalize);
    if (t + 4 <= S) {
        const float4 p = *reinterpret_cast<const float4 *>(x), q = *reinterpret_cast<const float4 *>(x + 4);
        const float l0[2] = {p.x, p.y}, l1[2] = {p.z, p.w}, l2[2] = {q.x, q.y}, l3[2] = {q.z, q.w};
        *reinterpret_cast<float4 *>(y) = make_float4(down_mix_sample(l0, 2, w), down_mix_sample(l1, 2, w),
                                                     down_mix_sample(l2, 2, w), down_mix_sample(l3, 2, w));
    } else {
        for (long long u = 0; u < S - t; u++) y[u] = down_mix_sample(x + 2 * u, 2, w);
    }
}

__global__ void __launch_bounds__(kThreads)
down_mix_kernel(const float *__restrict__ src, uint32_t srcStride, float *__restrict__ dst, uint32_t dstStride,
                const int *__restrict__ dims, int normalize, int skipStereo)
{
    const int b = blockIdx.y;
    const int C = dims[2 * b + 1];
    const long long t = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
    if (C <= 0 || (skipStereo && C == 2)) return;
    const long long S = min(static_cast<long long>(dims[2 * b]),
                            min(static_cast<long long>(srcStride) / C, static_cast<long long>(dstStride)));
    if (t >= S) return;
    const float *x = src + static_cast<size_t>(b) * srcStride + t * C;
    float *y = dst + static_cast<size_t>(b) * dstStride;
    y[t] = down_mix_sample(x, C, down_mix_weight(C, normalize));
}

}  // namespace
}  // namespace rppa

// Validation follows rppt_tensor_audio_augmentations.cpp:416-448.
RPPA_EXPORT RppStatus rppt_down_mixing_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                           RpptDescPtr dstDescPtr, Rpp32s *srcDimsTensor, bool normalizeWeights,
                                           rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    const size_t dims = srcDescPtr->numDims - 1;
    if (dims != 1 && dims != 2) return RPP_ERROR_INVALID_SRC_DIMS;
    if (srcDescPtr->dataType != F32 || dstDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcDimsTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    const unsigned B = srcDescPtr->n;
    if (B == 0) return RPP_SUCCESS;
    // 16-byte aligned rows take the stereo kernel for their two-channel clips (4 samples per thread)
    const bool vec = (reinterpret_cast<uintptr_t>(srcPtr) & 15) == 0 && (reinterpret_cast<uintptr_t>(dstPtr) & 15) == 0 &&
                     srcDescPtr->strides.nStride % 4 == 0 && dstDescPtr->strides.nStride % 4 == 0;
    if (vec) {
        dim3 g4(rppa::ceil_div(rppa::ceil_div(dstDescPtr->strides.nStride, 4u), rppa::kThreads), B);
        hipLaunchKernelGGL(rppa::down_mix_stereo4_kernel, g4, dim3(rppa::kThreads), 0, rppHandle->stream,
                           static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, static_cast<float *>(dstPtr),
                           dstDescPtr->strides.nStride, srcDimsTensor, normalizeWeights ? 1 : 0);
    }
    dim3 grid(rppa::ceil_div(dstDescPtr->strides.nStride, rppa::kThreads), B);
    hipLaunchKernelGGL(rppa::down_mix_kernel, grid, dim3(rppa::kThreads), 0, rppHandle->stream,
                       static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, static_cast<float *>(dstPtr),
                       dstDescPtr->strides.nStride, srcDimsTensor, normalizeWeights ? 1 : 0, vec ? 1 : 0);
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}
