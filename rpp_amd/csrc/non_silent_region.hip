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

constexpr int kChunk = 32;          // samples per lane per prefetch step (three steps in flight)
constexpr int kVecChunk = 16;       // the same with 16-byte loads
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
    int vec;            // rows 16-byte aligned, nStride and R multiples of 4: 16-byte loads
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
// visit(p, mms[p], p < end) in order, branch-free: the recurrence runs on to the end of the last
// chunk (values past `end` are flagged invalid and never used).  The recurrence is sequential (its rounding is the HOST's), so the
// loads run ahead of it: kDepth chunks of kChunk samples are in flight while one is consumed (the
// three buffers rotate by name in a loop unrolled three times, so no values are copied).
template <class Visit>
__device__ __forceinline__ void mms_block(const float *x, int n, int start, int end, int W, float mf, Visit &&visit)
{
    float sum = 0.0f;
    {
        // fresh sum of the W-1 samples before the block (HOST :129-130), in order
        const int f0 = max(start - W + 1, 0);
        float b0[kChunk], b1[kChunk], b2[kChunk];
        auto load = [&](float (&v)[kChunk], int i0) {
#pragma unroll
            for (int u = 0; u < kChunk; u++) v[u] = ld(x, i0 + u, n);
        };
        auto add = [&](const float (&v)[kChunk], int i0) {
#pragma unroll
            for (int u = 0; u < kChunk; u++)
                {
                    float q = v[u] * v[u];
                    sum = sum + (i0 + u < start ? q : 0.0f);   // adding +0 keeps the sum
                }
        };
        if (f0 < start) {
            load(b0, f0);
            load(b1, f0 + kChunk);
            load(b2, f0 + 2 * kChunk);
            for (int i0 = f0; i0 < start; i0 += 3 * kChunk) {
                add(b0, i0);
                load(b0, i0 + 3 * kChunk);
                add(b1, i0 + kChunk);
                load(b1, i0 + 4 * kChunk);
                add(b2, i0 + 2 * kChunk);
                load(b2, i0 + 5 * kChunk);
            }
        }
    }
    float a0[kChunk], a1[kChunk], a2[kChunk], s0[kChunk], s1[kChunk], s2[kChunk];
    auto load2 = [&](float (&av)[kChunk], float (&sv)[kChunk], int p0) {
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            av[i] = ld(x, p0 + i, n);
            sv[i] = ld(x, p0 - W + 1 + i, n);
        }
    };
    auto run = [&](const float (&av)[kChunk], const float (&sv)[kChunk], int p0) {
#pragma unroll
        for (int i = 0; i < kChunk; i++) {
            const int p = p0 + i;
            float q = av[i] * av[i];
            sum = sum + q;
            visit(p, sum * mf, p < end);
            float r = sv[i] * sv[i];
            sum = sum - (p - W + 1 >= 0 ? r : 0.0f);   // x - 0 is x: the HOST's skipped subtraction
        }
    };
    load2(a0, s0, start);
    load2(a1, s1, start + kChunk);
    load2(a2, s2, start + 2 * kChunk);
    for (int p0 = start; p0 < end; p0 += 3 * kChunk) {
        run(a0, s0, p0);
        load2(a0, s0, p0 + 3 * kChunk);
        run(a1, s1, p0 + kChunk);
        load2(a1, s1, p0 + 4 * kChunk);
        run(a2, s2, p0 + 2 * kChunk);
        load2(a2, s2, p0 + 5 * kChunk);
    }
}

// The same recurrence on 16-byte loads, for rows that are 16-byte aligned with a block start that
// is a multiple of 4: the operations and their order are mms_block's, only the loads differ (the
// trailing samples x[p - W + 1] are elements M .. M+3 of two aligned groups, M = (start - W + 1) & 3).
template <int M, class Visit>
__device__ __forceinline__ void mms_main_vec(const float *x, int rowLen, int start, int end, int W, float mf,
                                             float &sum, Visit &visit)
{
    constexpr int G = kVecChunk / 4;                       // 16-byte groups per chunk
    const int tb0 = start - W + 1 - M;                     // a multiple of 4
    auto ld4 = [&](int i) { return *reinterpret_cast<const float4 *>(x + min(max(i, 0), rowLen - 4)); };
    auto load = [&](float4 (&A)[G], float4 (&T)[G], int p0) {
        const int t0 = tb0 + (p0 - start);
#pragma unroll
        for (int g = 0; g < G; g++) {
            A[g] = ld4(p0 + 4 * g);
            T[g] = ld4(t0 + 4 * g);
        }
    };
    auto run = [&](const float4 (&A)[G], const float4 (&T)[G], const float4 &Tn, int p0) {
#pragma unroll
        for (int g = 0; g < G; g++) {
            const float4 t1 = g + 1 < G ? T[g + 1] : Tn;
            const float a[4] = {A[g].x, A[g].y, A[g].z, A[g].w};
            const float t8[8] = {T[g].x, T[g].y, T[g].z, T[g].w, t1.x, t1.y, t1.z, t1.w};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int p = p0 + 4 * g + u;
                float q = a[u] * a[u];
                sum = sum + q;
                visit(p, sum * mf, p < end);
                float r = t8[M + u] * t8[M + u];
                sum = sum - (p - W + 1 >= 0 ? r : 0.0f);
            }
        }
    };
    float4 A0[G], A1[G], A2[G], T0[G], T1[G], T2[G];
    load(A0, T0, start);
    load(A1, T1, start + kVecChunk);
    load(A2, T2, start + 2 * kVecChunk);
    for (int p0 = start; p0 < end; p0 += 3 * kVecChunk) {
        run(A0, T0, T1[0], p0);
        load(A0, T0, p0 + 3 * kVecChunk);
        run(A1, T1, T2[0], p0 + kVecChunk);
        load(A1, T1, p0 + 4 * kVecChunk);
        run(A2, T2, T0[0], p0 + 2 * kVecChunk);
        load(A2, T2, p0 + 5 * kVecChunk);
    }
}

template <class Visit>
__device__ __forceinline__ void mms_block_vec(const float *x, int rowLen, int start, int end, int W, float mf,
                                              Visit &&visit)
{
    constexpr int G = kVecChunk / 4;
    float sum = 0.0f;
    {
        // fresh sum of the W-1 samples before the block, in order, from aligned 16-byte groups
        const int f0 = max(start - W + 1, 0);
        const int a0 = f0 & ~3;
        auto ld4 = [&](int i) { return *reinterpret_cast<const float4 *>(x + min(max(i, 0), rowLen - 4)); };
        auto load = [&](float4 (&v)[G], int i0) {
#pragma unroll
            for (int g = 0; g < G; g++) v[g] = ld4(i0 + 4 * g);
        };
        auto add = [&](const float4 (&v)[G], int i0) {
#pragma unroll
            for (int g = 0; g < G; g++) {
                const float e[4] = {v[g].x, v[g].y, v[g].z, v[g].w};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int i = i0 + 4 * g + u;
                    float q = e[u] * e[u];
                    sum = sum + ((i >= f0 && i < start) ? q : 0.0f);   // adding +0 keeps the sum
                }
            }
        };
        if (f0 < start) {
            float4 b0[G], b1[G], b2[G];
            load(b0, a0);
            load(b1, a0 + kVecChunk);
            load(b2, a0 + 2 * kVecChunk);
            for (int i0 = a0; i0 < start; i0 += 3 * kVecChunk) {
                add(b0, i0);
                load(b0, i0 + 3 * kVecChunk);
                add(b1, i0 + kVecChunk);
                load(b1, i0 + 4 * kVecChunk);
                add(b2, i0 + 2 * kVecChunk);
                load(b2, i0 + 5 * kVecChunk);
            }
        }
    }
    switch ((start - W + 1) & 3) {
    case 0: mms_main_vec<0>(x, rowLen, start, end, W, mf, sum, visit); break;
    case 1: mms_main_vec<1>(x, rowLen, start, end, W, mf, sum, visit); break;
    case 2: mms_main_vec<2>(x, rowLen, start, end, W, mf, sum, visit); break;
    default: mms_main_vec<3>(x, rowLen, start, end, W, mf, sum, visit); break;
    }
}

__global__ void __launch_bounds__(kPass1Threads) nsr_blocks_kernel(NsrArgs a, int B)
{
    const long long lane = static_cast<long long>(blockIdx.x) * kPass1Threads + threadIdx.x;
    if (lane >= static_cast<long long>(B) * a.maxBlocks) return;
    // consecutive lanes take the same block of consecutive clips: the lanes of a wave then read
    // rows a clip stride apart rather than blocks a reset interval apart (a power-of-two stride
    // that sends every lane's load to the same memory channel)
    const int b = static_cast<int>(lane % B);
    const int j = static_cast<int>(lane / B);
    const long long cell = static_cast<long long>(b) * a.maxBlocks + j;
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
    auto visit = [&](int p, float v, bool valid) {
        m = fmaxf(m, valid ? v : -INFINITY);
        const bool hit = valid && knownCut && v >= cut;
        first = (hit && first < 0) ? p : first;
        last = hit ? p : last;
    };
    if (a.vec) mms_block_vec(x, static_cast<int>(a.nStride), start, end, a.W, a.meanFactor, visit);
    else mms_block(x, L, start, end, a.W, a.meanFactor, visit);
    a.blockMax[cell] = m;
    if (knownCut) {
        a.blockFirst[cell] = first;
        a.blockLast[cell] = last;
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
            auto visit = [&](int p, float v, bool valid) {
                const bool h = valid && !done && v >= cut;
                hit = h ? p : hit;
                done = done || (h && head);  // the head wants the first hit; the tail keeps the last
            };
            const int s0 = j * R, e0 = static_cast<int>(min(static_cast<long long>(L), static_cast<long long>(j) * R + R));
            if (a.vec) mms_block_vec(x, static_cast<int>(a.nStride), s0, e0, a.W, a.meanFactor, visit);
            else mms_block(x, L, s0, e0, a.W, a.meanFactor, visit);
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
    a.vec = (reinterpret_cast<uintptr_t>(srcPtr) & 15) == 0 && nStride % 4 == 0 && (resetInterval <= 0 || resetInterval % 4 == 0) ? 1 : 0;
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
