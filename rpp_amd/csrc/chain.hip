// chain.hip -- rppt_logmel_chain_gpu: the reference's four-call chain behind one entry point.
//
// pre-emphasis -> spectrogram (NFT, power) -> mel filter bank -> to_decibels through the drop-in
// entry points, with the three intermediates in the handle's workspace.  perClip != 0 issues the
// four calls once per clip with single-clip descriptors (one launch per clip per stage -- the
// plain per-clip GPU path); perClip == 0 issues them once for the whole batch, as a caller of
// the reference does (no in-tree composition exists; utilities/test_suite/HIP/Tensor_audio_hip.cpp
// drives the stages one by one).  srcLengthTensor must be host-readable (host or pinned memory).
#include "common.hpp"

#include <cstring>

RPPA_EXPORT RppStatus rppt_logmel_chain_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                            RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor,
                                            const RpptLogMelParams *p, int perClip, rppHandle_t h)
{
    if (!h || !srcDescPtr || !dstDescPtr || !p || !srcPtr || !dstPtr || !srcLengthTensor)
        return RPP_ERROR_INVALID_ARGUMENTS;
    if (dstDescPtr->layout != NFT) return RPP_ERROR_INVALID_DST_LAYOUT;
    const int B = static_cast<int>(srcDescPtr->n);
    if (B == 0) return RPP_SUCCESS;
    const int nfft = p->nfft ? p->nfft : p->windowLength;
    const int K = nfft / 2 + 1, M = p->numFilter;
    const uint32_t hs = dstDescPtr->strides.hStride;
    if (hs == 0 || dstDescPtr->h < static_cast<unsigned>(M)) return RPP_ERROR_INVALID_DST_DIMS;

    // small per-clip arrays, staged through pinned memory into the device workspace
    const size_t smallInts = static_cast<size_t>(B) * 6 + 64;  // lengths, mel dims, dB dims, coeffs
    const size_t nS = srcDescPtr->strides.nStride;
    const size_t peFloats = static_cast<size_t>(B) * nS;
    const size_t specFloats = static_cast<size_t>(B) * K * hs;
    const size_t melFloats = static_cast<size_t>(B) * M * hs;
    const size_t bytes = (smallInts + peFloats + specFloats + melFloats) * 4 + 256;
    char *ws = static_cast<char *>(rppa::workspace(h, bytes));
    if (!ws) return RPP_ERROR_NOT_ENOUGH_MEMORY;
    if (h->pinnedBytes < smallInts * 4) {
        if (h->pinned) (void)hipHostFree(h->pinned);
        h->pinned = nullptr;
        h->pinnedBytes = 0;
        RPPA_HIP(hipHostMalloc(&h->pinned, smallInts * 4, hipHostMallocDefault));
        h->pinnedBytes = smallInts * 4;
    }
    RPPA_HIP(hipStreamSynchronize(h->stream));  // the staging buffer may still feed a previous copy
    int *st = static_cast<int *>(h->pinned);
    int *sLen = st, *sMel = st + B, *sDb = st + 3 * B;
    float *sCoeff = reinterpret_cast<float *>(st + 5 * B);
    for (int b = 0; b < B; b++) {
        int L = std::min(srcLengthTensor[b], static_cast<int>(nS));
        int F = rppa::num_windows(std::max(L, 0), p->windowLength, p->windowStep, p->centerWindows != 0);
        F = std::max(0, std::min(F, static_cast<int>(hs)));
        sLen[b] = L;
        sMel[2 * b] = K;
        sMel[2 * b + 1] = F;
        sDb[2 * b] = F;      // RpptImagePatch.width  = frames
        sDb[2 * b + 1] = M;  // RpptImagePatch.height = mel rows
        sCoeff[b] = p->preEmphasisCoeff;
    }
    int *dSmall = reinterpret_cast<int *>(ws);
    RPPA_HIP(hipMemcpyAsync(dSmall, st, static_cast<size_t>(B) * 6 * 4, hipMemcpyHostToDevice, h->stream));
    int *dLen = dSmall, *dMel = dSmall + B, *dDb = dSmall + 3 * B;
    float *dCoeff = reinterpret_cast<float *>(dSmall + 5 * B);
    float *pe = reinterpret_cast<float *>(ws + smallInts * 4);
    float *spec = pe + peFloats;
    float *mel = spec + specFloats;

    auto run = [&](int b0, int n) -> RppStatus {
        RpptDesc sd = *srcDescPtr;
        sd.n = n;
        RpptDesc pd = sd;
        RpptDesc specd{};
        specd.numDims = 3;
        specd.dataType = F32;
        specd.n = n; specd.c = 1; specd.h = K; specd.w = hs;
        specd.strides = {static_cast<Rpp32u>(K * hs), 1, hs, 1};
        specd.layout = NFT;
        RpptDesc meld = specd;
        meld.h = M;
        meld.strides.nStride = M * hs;
        RpptDesc dbd = *dstDescPtr;
        dbd.n = n;
        dbd.numDims = 3;
        const float *s = static_cast<const float *>(srcPtr) + static_cast<size_t>(b0) * nS;
        float *peb = pe + static_cast<size_t>(b0) * nS;
        float *specb = spec + static_cast<size_t>(b0) * K * hs;
        float *melb = mel + static_cast<size_t>(b0) * M * hs;
        float *dstb = static_cast<float *>(dstPtr) + static_cast<size_t>(b0) * dstDescPtr->strides.nStride;
        RppStatus r;
        r = rppt_pre_emphasis_filter_gpu(const_cast<float *>(s), &sd, peb, &pd, dLen + b0, dCoeff + b0,
                                         p->borderType, h);
        if (r != RPP_SUCCESS) return r;
        r = rppt_spectrogram_gpu(peb, &pd, specb, &specd, dLen + b0, p->centerWindows != 0, p->reflectPadding != 0,
                                 nullptr, nfft, p->power, p->windowLength, p->windowStep, h);
        if (r != RPP_SUCCESS) return r;
        r = rppt_mel_filter_bank_gpu(specb, &specd, melb, &meld, dMel + 2 * b0, p->maxFreq, p->minFreq,
                                     p->melFormula, M, p->sampleRate, p->normalize != 0, h);
        if (r != RPP_SUCCESS) return r;
        return rppt_to_decibels_gpu(melb, &meld, dstb, &dbd, reinterpret_cast<RpptImagePatch *>(dDb + 2 * b0),
                                    p->cutOffDB, p->multiplier, p->referenceMagnitude, h);
    };
    if (perClip) {
        for (int b = 0; b < B; b++) {
            RppStatus r = run(b, 1);
            if (r != RPP_SUCCESS) return r;
        }
        return RPP_SUCCESS;
    }
    return run(0, B);
}
