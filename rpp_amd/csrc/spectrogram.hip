// spectrogram.hip -- rppt_spectrogram_gpu (api/rppt_tensor_audio_augmentations.h:232), FP32.
//
// Semantics of cpu/kernel/spectrogram.cpp:71-251: F = L/H + 1 centered frames (else (L-W)/H + 1),
// frame f starts at f*H - (center ? W/2 : 0), out-of-range samples are mirrored without
// repeating the edge (reflect) or zero, the frame is multiplied by the window, and bin
// k in [0, N/2] gets |X_k|^2 when power == 2 and |X_k| for any other value -- the rule of the
// GPU entry point this replaces (hip/kernel/spectrogram.cpp:231-232; the HOST kernel instead
// takes |X_k| only for power == 1, cpu/kernel/spectrogram.cpp:217-242).  NFT writes
// dst[k*hStride + f], NTF writes dst[f*hStride + k]; only frames < F are written.
//
// The DFT is a dense FP32 contraction  C[f][c] = sum_t x[f*H - off + t] * basis[t][c]  against
// the window-folded basis of tables.hip (the window multiply happens once, in FP64, when the
// basis is built).  The framing is implicit: the A tile is gathered straight from the signal
// with the reflect mapping, so no frame matrix ever goes to HBM (the reference writes
// B*maxFrames*nfft floats of frames and reads them back, hip/kernel/spectrogram.cpp:100-143).
//
// Tiling (one block = 32 frames of one clip, 256 threads = 4 waves):
//   per K step of 16 samples the block stages A[16][32] and basis[16][512] in LDS; each thread
//   owns 4 frames x 8 re/im column pairs (pairs cg, cg + 32, ... so consecutive lanes read
//   consecutive LDS words) of the 32 x 512 output tile in registers;
//   power is formed in registers, staged in LDS and stored coalesced along frames (NFT) or
//   bins (NTF).  Columns beyond 512 (nfft > 512) are processed in further 512-column passes.
#include "common.hpp"

namespace rppa {
namespace {

constexpr int kFT = 32;      // frames per block
constexpr int kKC = 16;      // K step
constexpr int kCC = 512;     // columns per pass
constexpr int kThreads = 256;
constexpr int kStage = (kKC * kCC > kFT * (kCC / 2 + 1)) ? kKC * kCC : kFT * (kCC / 2 + 1);

struct SpecArgs {
    const float *src;
    uint32_t srcStride;
    const int *lengths;
    float *dst;
    uint32_t dstNStride, hStride;
    int nft;
    int center, reflect;
    int N, W, H, power, cols;
    const float *basis;  // [W][cols]
};

__device__ inline float sample_at(const float *x, int L, int i, int reflect)
{
    if (i >= 0 && i < L) return x[i];
    if (L <= 0) return 0.0f;
    if (reflect) return x[idx_reflect(i, L)];
    return 0.0f;
}

__global__ void __launch_bounds__(kThreads)
spectrogram_fp32_kernel(SpecArgs a)
{
    __shared__ float sA[kKC][kFT];          // A chunk, [t][f]
    __shared__ float sB[kStage];            // basis chunk [t][c]; reused as the power stage
    const int b = blockIdx.y;
    const int L = min(a.lengths[b], static_cast<int>(a.srcStride));  // never read past the clip's row
    const int F = num_windows(L, a.W, a.H, a.center != 0);
    const int f0 = blockIdx.x * kFT;
    if (f0 >= F) return;
    const float *x = a.src + static_cast<size_t>(b) * a.srcStride;
    const int off = a.center ? a.W / 2 : 0;
    const int tid = threadIdx.x;
    const int cg = tid & 31;        // column pairs cg, cg + 32, ..., cg + 224 of a pass
    const int fg = tid >> 5;        // frame group: 4 frames
    const int K = a.N / 2 + 1;
    const int nPasses = (a.cols + kCC - 1) / kCC;
    float *stage = sB;              // [kFT][kCC/2 + 1] power stage after the K loop

    for (int pass = 0; pass < nPasses; pass++) {
        const int cBase = pass * kCC;
        float acc[4][16];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 16; j++) acc[i][j] = 0.0f;

        for (int t0 = 0; t0 < a.W; t0 += kKC) {
            // stage A: 16 x 32 = 512 values, 2 per thread
            for (int e = tid; e < kKC * kFT; e += kThreads) {
                int tt = e / kFT, ff = e % kFT;
                int t = t0 + tt, f = f0 + ff;
                float v = 0.0f;
                if (t < a.W && f < F) v = sample_at(x, L, f * a.H - off + t, a.reflect);
                sA[tt][ff] = v;
            }
            // stage basis rows t0..t0+15, columns cBase..cBase+511
            for (int e = tid; e < kKC * kCC; e += kThreads) {
                int tt = e / kCC, cc = e % kCC;
                int t = t0 + tt, c = cBase + cc;
                sB[e] = (t < a.W && c < a.cols) ? a.basis[static_cast<size_t>(t) * a.cols + c] : 0.0f;
            }
            __syncthreads();
#pragma unroll 4
            for (int kk = 0; kk < kKC; kk++) {
                float av[4];
#pragma unroll
                for (int i = 0; i < 4; i++) av[i] = sA[kk][fg * 4 + i];
                float bv[16];
#pragma unroll
                for (int j = 0; j < 8; j++) {   // column pair cg + 32 j: consecutive lanes, consecutive pairs
                    const float2 b2 = *reinterpret_cast<const float2 *>(&sB[kk * kCC + 2 * (cg + 32 * j)]);
                    bv[2 * j] = b2.x;
                    bv[2 * j + 1] = b2.y;
                }
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 16; j++) acc[i][j] = fmaf(av[i], bv[j], acc[i][j]);
            }
            __syncthreads();
        }

        // power into the stage: stage[f][p] for pair p of this pass (pair 0 of pass 0 = DC/Nyquist)
        const int pairsInPass = kCC / 2;
#pragma unroll
        for (int i = 0; i < 4; i++) {
#pragma unroll
            for (int j = 0; j < 8; j++) {
                float re = acc[i][2 * j], im = acc[i][2 * j + 1];
                int p = cg + 32 * j;
                float v;
                float extra = -1.0f;
                if (pass == 0 && p == 0) {
                    v = re * re;            // DC
                    extra = im * im;        // Nyquist (column 1)
                } else {
                    v = re * re + im * im;
                }
                if (a.power == 1) {
                    v = sqrtf(v);
                    if (extra >= 0.0f) extra = sqrtf(extra);
                }
                stage[(fg * 4 + i) * (pairsInPass + 1) + p] = v;
                if (pass == 0 && p == 0) stage[(fg * 4 + i) * (pairsInPass + 1) + pairsInPass] = extra;
            }
        }
        __syncthreads();
        // store: bins of this pass are pass*256 + p (p >= 1, or p == 0 for pass > 0); pass 0 also
        // owns bin 0 (slot 0) and, for even N, bin N/2 (slot 256)
        const int nf = min(kFT, F - f0);
        const int pStart = pass * pairsInPass;
        for (int e = tid; e < (pairsInPass + 1) * kFT; e += kThreads) {
            int slot, f;
            if (a.nft) { slot = e / kFT; f = e % kFT; } else { f = e / (pairsInPass + 1); slot = e % (pairsInPass + 1); }
            if (f >= nf) continue;
            int k;
            if (slot == pairsInPass) {
                if (pass != 0 || (a.N & 1)) continue;
                k = a.N / 2;
            } else {
                k = pStart + slot;
                if (k >= K) continue;
                if (k == a.N / 2 && !(a.N & 1) && k != 0) continue;  // Nyquist comes from column 1
            }
            float v = stage[f * (pairsInPass + 1) + slot];
            size_t o = static_cast<size_t>(b) * a.dstNStride +
                       (a.nft ? static_cast<size_t>(k) * a.hStride + (f0 + f) : static_cast<size_t>(f0 + f) * a.hStride + k);
            a.dst[o] = v;
        }
        __syncthreads();
    }
}

}  // namespace
}  // namespace rppa

// Validation follows rppt_tensor_audio_augmentations.cpp:579-629 (layout NFT/NTF, src dims 1,
// dst dims 2, F32).  W > nfft is rejected (the reference writes past its FFT buffer there).
RPPA_EXPORT RppStatus rppt_spectrogram_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                           RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor, bool centerWindows,
                                           bool reflectPadding, Rpp32f *windowFunction, Rpp32s nfft, Rpp32s power,
                                           Rpp32s windowLength, Rpp32s windowStep, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (dstDescPtr->layout != NFT && dstDescPtr->layout != NTF) return RPP_ERROR_INVALID_DST_LAYOUT;
    if (srcDescPtr->numDims - 1 != 1) return RPP_ERROR_INVALID_SRC_DIMS;
    if (dstDescPtr->numDims - 1 != 2) return RPP_ERROR_INVALID_DST_DIMS;
    if (srcDescPtr->dataType != F32 || dstDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcLengthTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    if (nfft == 0) nfft = windowLength;
    if (windowLength <= 0 || windowStep <= 0 || nfft < 2 || windowLength > nfft) return RPP_ERROR_INVALID_ARGUMENTS;
    const bool nft = dstDescPtr->layout == NFT;
    const int K = nfft / 2 + 1;
    const unsigned maxFrames = nft ? dstDescPtr->w : dstDescPtr->h;
    const unsigned B = srcDescPtr->n;
    if (B == 0 || maxFrames == 0) return RPP_SUCCESS;
    if (nft ? (dstDescPtr->strides.hStride < maxFrames) : (dstDescPtr->strides.hStride < static_cast<unsigned>(K)))
        return RPP_ERROR_INVALID_DST_DIMS;

    const rppa::DftTables *t = rppa::get_dft_tables(rppHandle, nfft, windowLength, windowFunction, -1);
    if (!t) return RPP_ERROR;
    rppa::SpecArgs a;
    a.src = static_cast<const float *>(srcPtr);
    a.srcStride = srcDescPtr->strides.nStride;
    a.lengths = srcLengthTensor;
    a.dst = static_cast<float *>(dstPtr);
    a.dstNStride = dstDescPtr->strides.nStride;
    a.hStride = dstDescPtr->strides.hStride;
    a.nft = nft;
    a.center = centerWindows;
    a.reflect = reflectPadding;
    a.N = nfft;
    a.W = windowLength;
    a.H = windowStep;
    a.power = power == 2 ? 2 : 1;   // the kernel's flag: 1 = magnitude, 2 = squared magnitude
    a.cols = t->cols;
    a.basis = t->basis;
    dim3 grid(rppa::ceil_div(maxFrames, rppa::kFT), B);
    hipLaunchKernelGGL(rppa::spectrogram_fp32_kernel, grid, dim3(rppa::kThreads), 0, rppHandle->stream, a);
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}
