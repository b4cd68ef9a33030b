// to_decibels.hip -- rppt_to_decibels_gpu (api/rppt_tensor_audio_augmentations.h:122).
//
// out = (multiplier * log10(2)) * log2(max(minRatio, x * (1/ref)))  with
// minRatio = 10^(cutOffDB/multiplier) (0 -> smallest subnormal) and ref = referenceMagnitude, or
// the clip's maximum when referenceMagnitude == 0 (cpu/kernel/to_decibels.cpp:27-99).
// srcDims[b] = {width = columns, height = rows}; width == 1 means 1-D data of `height` samples.
// log2 is evaluated in FP64 and rounded once, the correctly rounded FP32 log2 the HOST path's
// log2f returns, so outputs are bit-identical to RPP HOST.  No host synchronisation: the
// per-clip maximum, when needed, is an extra reduction kernel on the same stream (the reference
// synchronises twice, hip/kernel/to_decibels.cpp:283, :295).
#include "common.hpp"

#include <cfloat>
#include <cmath>

namespace rppa {
namespace {

constexpr int kThreads = 256;

__device__ inline int float_order(float v)
{
    int i = __float_as_int(v);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ inline float order_float(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__global__ void max_init_kernel(int *maxEnc, int B)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) maxEnc[b] = float_order(-FLT_MAX);
}

__global__ void __launch_bounds__(kThreads)
clip_max_kernel(const float *__restrict__ src, uint32_t nStride, uint32_t hStride, const RpptImagePatch *dims, int *maxEnc)
{
    const int b = blockIdx.z;
    const int row = blockIdx.y;
    const unsigned width = dims[b].width, height = dims[b].height;
    const bool oneD = width == 1;
    const unsigned cols = oneD ? min(height, nStride) : min(width, hStride);   // stay inside the clip's buffer
    if (!oneD && static_cast<unsigned>(row) >= height) return;
    if (oneD && row != 0) return;
    const float *x = src + static_cast<size_t>(b) * nStride + (oneD ? 0 : static_cast<size_t>(row) * hStride);
    float m = -FLT_MAX;
    for (unsigned c = blockIdx.x * kThreads + threadIdx.x; c < cols; c += gridDim.x * kThreads) m = fmaxf(m, x[c]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    __shared__ float red[kThreads / kWave];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = red[0];
        for (int i = 1; i < kThreads / kWave; i++) r = fmaxf(r, red[i]);
        atomicMax(&maxEnc[b], float_order(r));
    }
}

__device__ inline float db_value(float x, float inv, float minRatio, float mult)
{
    float v = x * inv;
    float m = (minRatio < v) ? v : minRatio;   // std::max(minRatio, v)
    return mult * static_cast<float>(log2(static_cast<double>(m)));
}

__global__ void __launch_bounds__(kThreads)
to_db_kernel(const float *__restrict__ src, uint32_t srcNStride, uint32_t srcHStride, float *__restrict__ dst,
             uint32_t dstNStride, uint32_t dstHStride, const RpptImagePatch *dims, float minRatio, float mult,
             float refMag, const int *maxEnc)
{
    const int b = blockIdx.z;
    const int row = blockIdx.y;
    const unsigned width = dims[b].width, height = dims[b].height;
    const bool oneD = width == 1;
    const unsigned cols = oneD ? min(height, min(srcNStride, dstNStride)) : min(width, min(srcHStride, dstHStride));
    if (!oneD && static_cast<unsigned>(row) >= height) return;
    if (oneD && row != 0) return;
    float ref = maxEnc ? order_float(maxEnc[b]) : refMag;
    float inv = (ref != 0.0f) ? (1.f / ref) : 1.0f;
    const float *x = src + static_cast<size_t>(b) * srcNStride + (oneD ? 0 : static_cast<size_t>(row) * srcHStride);
    float *y = dst + static_cast<size_t>(b) * dstNStride + (oneD ? 0 : static_cast<size_t>(row) * dstHStride);
    for (unsigned c = blockIdx.x * kThreads + threadIdx.x; c < cols; c += gridDim.x * kThreads)
        y[c] = db_value(x[c], inv, minRatio, mult);
}

}  // namespace
}  // namespace rppa

// Validation follows rppt_tensor_audio_augmentations.cpp:373-412.
RPPA_EXPORT RppStatus rppt_to_decibels_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                           RpptDescPtr dstDescPtr, RpptImagePatchPtr srcDims, Rpp32f cutOffDB,
                                           Rpp32f multiplier, Rpp32f referenceMagnitude, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    const size_t dims = srcDescPtr->numDims - 1;
    if (dims != 1 && dims != 2) return RPP_ERROR_INVALID_SRC_DIMS;
    if (!multiplier) return RPP_ERROR_ZERO_DIVISION;
    if (srcDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcDims) return RPP_ERROR_INVALID_ARGUMENTS;
    const unsigned B = srcDescPtr->n;
    if (B == 0) return RPP_SUCCESS;

    // host-side constants exactly as cpu/kernel/to_decibels.cpp:38-43 computes them
    float minRatio = static_cast<float>(std::pow(10.0, static_cast<double>(cutOffDB / multiplier)));
    if (minRatio == 0.0f) minRatio = std::nextafter(0.0f, 1.0f);
    const float log10Factor = 0.3010299956639812f;
    const float mult = multiplier * log10Factor;

    // grid: x over columns, y over rows (1 for 1-D data), z over clips
    const bool oneD = dims == 1;
    const unsigned cols = oneD ? srcDescPtr->strides.nStride : srcDescPtr->strides.hStride;
    const unsigned rows = oneD ? 1 : srcDescPtr->h;
    const unsigned gx = std::min(rppa::ceil_div(cols, rppa::kThreads), 4096u);
    dim3 grid(gx ? gx : 1, rows ? rows : 1, B);
    int *maxEnc = nullptr;
    if (referenceMagnitude == 0.0f) {
        maxEnc = static_cast<int *>(rppa::workspace(rppHandle, sizeof(int) * B));
        if (!maxEnc) return RPP_ERROR_NOT_ENOUGH_MEMORY;
        hipLaunchKernelGGL(rppa::max_init_kernel, dim3(rppa::ceil_div(B, 256)), dim3(256), 0, rppHandle->stream, maxEnc,
                           static_cast<int>(B));
        hipLaunchKernelGGL(rppa::clip_max_kernel, grid, dim3(rppa::kThreads), 0, rppHandle->stream,
                           static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, srcDescPtr->strides.hStride,
                           srcDims, maxEnc);
    }
    hipLaunchKernelGGL(rppa::to_db_kernel, grid, dim3(rppa::kThreads), 0, rppHandle->stream,
                       static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, srcDescPtr->strides.hStride,
                       static_cast<float *>(dstPtr), dstDescPtr->strides.nStride, dstDescPtr->strides.hStride, srcDims,
                       minRatio, mult, referenceMagnitude, maxEnc);
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}
