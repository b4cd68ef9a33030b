// non_silent_region.hip -- rppt_non_silent_region_detection_gpu
// (api/rppt_tensor_audio_augmentations.h:86).
//
// Semantics of the HOST kernel (src/modules/tensor/cpu/kernel/non_silent_region_detection.cpp:93-191):
// the moving mean square mms[p] = (sum of x[i]^2, p-W < i <= p) / W is a running FP32 sum that
// is restarted every resetInterval outputs (from a fresh sum of the W-1 samples before the
// block); the region runs from the first to the last p with mms[p] >= base * 10^(cutOffDB/10),
// base = max(mms) (referencePower == 0) or referencePower, and its begin is moved W-1 samples
// earlier (clamped at 0) unless it is 0.
//
// The running sum is evaluated in the HOST order (add the square, scale, subtract the square
// leaving the window; products rounded before the add, as the reference's -mno-fma build does),
// so mms is bit-identical to that sequence and the detected indices match it exactly.  The
// reference's HIP kernel uses per-tile prefix sums instead, whose rounding differs from HOST.
//
// Parallelism is the reset blocks: one lane runs one block's recurrence.  Pass 1 gives each
// block's maximum (and, when the cut-off is known up front, its first and last hit); pass 2, one
// workgroup per clip, reduces them and rescans only the two blocks holding the region's ends.
#include <cmath>

#include "common.hpp"

namespace rppa {
namespace {

constexpr int kChunk = 16;          // samples per lane per prefetch step
constexpr int kPass1Threads = 64;   // one wave per workgroup spreads the few lanes over CUs
constexpr int kPass2Threads = 128;

struct NsrArgs {
    const float *src;
    long long nStride;
    const int *len;
    int W;
    float meanFactor;
    int R;              // reset length; <= 0 means the whole clip is one block
    int maxBlocks;
    float cutOff;       // 10^(cutOffDB/10)
    float refPower;     // 0: per-clip maximum
    float *blockMax;    // [B][maxBlocks]
    int *blockFirst;    // [B][maxBlocks], used when refPower != 0
    int *blockLast;
    int *begin;
    int *length;
};

__device__ __forceinline__ int clip_len(const NsrArgs &a, int b)
{
    long long L = a.len[b];
    if (L < 0) L = 0;
    if (L > a.nStride) L = a.nStride;
    return static_cast<int>(L);
}

__device__ __forceinline__ int blocks_of(int L, int R) { return L <= 0 ? 0 : (R <= 0 ? 1 : (L + R - 1) / R); }

// x[i] for i clamped into [0, n): chunk loads may run past either end of a stream; the values
// loaded there are never used.
__device__ __forceinline__ float ld(const float *x, int i, int n) { return x[min(max(i, 0), n - 1)]; }

// Runs the HOST recurrence over outputs [start, end) of one clip of length n and calls
// visit(p, mms[p]) in order.
template <class Visit>
__device__ __forceinline__ void mms_block(const float *x, int n, int start, int end, int W, float mf, Visit &&visit)
{
    float sum = 0.0f;
    // fresh sum of the W-1 samples before the block (HOST :129-130)
    for (int i = max(start - W + 1, 0); i < start; i++) {
        float q = x[i] * x[i];
        sum = sum + q;
    }
    float ca[kChunk], cs[kChunk];
#pragma unroll
    for (int i = 0; i < kChunk; i++) {
        ca[i] = ld(x, start + i, n);
        cs[i] = ld(x, start - W + 1 + i, n);
    }
    for (int p0 = start; p0 < end; p0 += kChunk) {
        float na[kChunk], ns[kChunk];
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            na[i] = ld(x, p0 + kChunk + i, n);
            ns[i] = ld(x, p0 + kChunk - W + 1 + i, n);
        }
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            const int p = p0 + i;
            if (p < end) {
                float q = ca[i] * ca[i];
                sum = sum + q;
                visit(p, sum * mf);
                if (p - W + 1 >= 0) {
                    float r = cs[i] * cs[i];
                    sum = sum - r;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            ca[i] = na[i];
            cs[i] = ns[i];
        }
    }
}

__global__ void __launch_bounds__(kPass1Threads) nsr_blocks_kernel(NsrArgs a, int B)
{
    const long long lane = static_cast<long long>(blockIdx.x) * kPass1Threads + threadIdx.x;
    if (lane >= static_cast<long long>(B) * a.maxBlocks) return;
    const int b = static_cast<int>(lane / a.maxBlocks);
    const int j = static_cast<int>(lane % a.maxBlocks);
    const int L = clip_len(a, b);
    if (j >= blocks_of(L, a.R)) return;
    const int R = a.R <= 0 ? L : a.R;
    const int start = j * R;
    const int end = static_cast<int>(min(static_cast<long long>(L), static_cast<long long>(start) + R));
    const float *x = a.src + b * a.nStride;
    const bool knownCut = a.refPower != 0.0f;
    const float cut = a.refPower * a.cutOff;
    float m = -INFINITY;
    int first = -1, last = -1;
    mms_block(x, L, start, end, a.W, a.meanFactor, [&](int p, float v) {
        m = fmaxf(m, v);
        if (knownCut && v >= cut) {
            if (first < 0) first = p;
            last = p;
        }
    });
    a.blockMax[lane] = m;
    if (knownCut) {
        a.blockFirst[lane] = first;
        a.blockLast[lane] = last;
    }
}

__global__ void __launch_bounds__(kPass2Threads) nsr_region_kernel(NsrArgs a)
{
    const int b = blockIdx.x;
    const int L = clip_len(a, b);
    const int nb = blocks_of(L, a.R);
    const bool knownCut = a.refPower != 0.0f;
    __shared__ float sMax[kPass2Threads];
    __shared__ int sFirstBlk, sLastBlk, sBegin, sEnd;
    const float *bm = a.blockMax + static_cast<long long>(b) * a.maxBlocks;

    float m = -INFINITY;
    for (int j = threadIdx.x; j < nb; j += kPass2Threads) m = fmaxf(m, bm[j]);
    sMax[threadIdx.x] = m;
    if (threadIdx.x == 0) {
        sFirstBlk = nb;
        sLastBlk = -1;
        sBegin = L;
        sEnd = -1;
    }
    __syncthreads();
    for (int s = kPass2Threads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sMax[threadIdx.x] = fmaxf(sMax[threadIdx.x], sMax[threadIdx.x + s]);
        __syncthreads();
    }
    const float base = knownCut ? a.refPower : sMax[0];
    const float cut = base * a.cutOff;

    if (knownCut) {
        // pass 1 already found every block's first and last hit
        const int *bf = a.blockFirst + static_cast<long long>(b) * a.maxBlocks;
        const int *bl = a.blockLast + static_cast<long long>(b) * a.maxBlocks;
        for (int j = threadIdx.x; j < nb; j += kPass2Threads) {
            if (bf[j] >= 0) {
                atomicMin(&sBegin, bf[j]);
                atomicMax(&sEnd, bl[j]);
            }
        }
    } else {
        for (int j = threadIdx.x; j < nb; j += kPass2Threads) {
            if (bm[j] >= cut) {
                atomicMin(&sFirstBlk, j);
                atomicMax(&sLastBlk, j);
            }
        }
        __syncthreads();
        // wave 0 lane 0 rescans the first block holding a hit, wave 1 lane 0 the last one
        const int R = a.R <= 0 ? L : a.R;
        const float *x = a.src + b * a.nStride;
        if ((threadIdx.x == 0 || threadIdx.x == kWave) && sLastBlk >= 0) {
            const bool head = threadIdx.x == 0;
            const int j = head ? sFirstBlk : sLastBlk;
            int hit = -1;
            bool done = false;
            mms_block(x, L, j * R, static_cast<int>(min(static_cast<long long>(L), static_cast<long long>(j) * R + R)), a.W, a.meanFactor, [&](int p, float v) {
                if (!done && v >= cut) {
                    hit = p;
                    done = head;  // the head wants the first hit; the tail keeps the last
                }
            });
            if (head) sBegin = hit;
            else sEnd = hit;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int begin = 0, length = 0;
        if (L > 0 && sEnd >= 0 && sBegin <= sEnd) {
            begin = sBegin;
            length = sEnd - sBegin + 1;
            if (begin != 0) {  // HOST :179-184
                const int nbgn = max(begin - (a.W - 1), 0);
                length += begin - nbgn;
                begin = nbgn;
            }
        }
        a.begin[b] = begin;
        a.length[b] = length;
    }
}

}  // namespace
}  // namespace rppa

// Validation follows rppt_tensor_audio_augmentations.cpp:332-369; windowLength < 1 and a
// resetInterval of 0 or below -1 (which make the HOST loop divide by zero or never advance) are
// rejected with RPP_ERROR_INVALID_ARGUMENTS.
RPPA_EXPORT RppStatus rppt_non_silent_region_detection_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr,
                                                           Rpp32s *srcLengthTensor, Rpp32s *detectedIndexTensor,
                                                           Rpp32s *detectionLengthTensor, Rpp32f cutOffDB,
                                                           Rpp32s windowLength, Rpp32f referencePower,
                                                           Rpp32s resetInterval, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (srcDescPtr->numDims - 1 != 1) return RPP_ERROR_INVALID_SRC_DIMS;
    if (srcDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !srcLengthTensor || !detectedIndexTensor || !detectionLengthTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    if (windowLength < 1 || resetInterval == 0 || resetInterval < -1) return RPP_ERROR_INVALID_ARGUMENTS;
    const int B = static_cast<int>(srcDescPtr->n);
    if (B == 0) return RPP_SUCCESS;
    const long long nStride = srcDescPtr->strides.nStride;
    if (nStride < 1 || nStride > 0x7fffffffLL) return RPP_ERROR_INVALID_ARGUMENTS;

    rppa::NsrArgs a{};
    a.src = static_cast<const float *>(srcPtr);
    a.nStride = nStride;
    a.len = srcLengthTensor;
    a.W = windowLength;
    a.meanFactor = 1.0f / static_cast<float>(windowLength);
    a.R = resetInterval;
    a.maxBlocks = resetInterval <= 0 ? 1 : static_cast<int>((nStride + resetInterval - 1) / resetInterval);
    a.cutOff = std::pow(10.0f, cutOffDB * 0.1f);
    a.refPower = referencePower;
    a.begin = detectedIndexTensor;
    a.length = detectionLengthTensor;
    const size_t cells = static_cast<size_t>(B) * a.maxBlocks;
    char *ws = static_cast<char *>(rppa::workspace(rppHandle, cells * 12));
    if (!ws) return RPP_ERROR_NOT_ENOUGH_MEMORY;
    a.blockMax = reinterpret_cast<float *>(ws);
    a.blockFirst = reinterpret_cast<int *>(ws + cells * 4);
    a.blockLast = reinterpret_cast<int *>(ws + cells * 8);

    const unsigned g1 = static_cast<unsigned>((cells + rppa::kPass1Threads - 1) / rppa::kPass1Threads);
    hipLaunchKernelGGL(rppa::nsr_blocks_kernel, dim3(g1), dim3(rppa::kPass1Threads), 0, rppHandle->stream, a, B);
    RPPA_HIP(hipGetLastError());
    hipLaunchKernelGGL(rppa::nsr_region_kernel, dim3(B), dim3(rppa::kPass2Threads), 0, rppHandle->stream, a);
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}
