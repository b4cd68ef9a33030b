// ref_shim.cpp -- C entry points onto RPP's own HOST audio kernels, compiled from the
// reference sources into oracle/_ref/libref_audio.so.  TEST INFRASTRUCTURE ONLY: the
// tests use it to check the oracle's restatement; nothing in the product links it.
//
// Each entry point builds the descriptors the reference test driver builds
// (utilities/test_suite/rpp_test_suite_audio.h:84-115), creates a HOST handle
// (src/modules/handle_api.cpp:29-44) and calls the *_host_tensor executor
// (src/include/tensor/host_tensor_executors.hpp:163-256).
#include "host_tensor_executors.hpp"
#include "rpp.h"

namespace {
RpptDesc desc_1d(int B, unsigned stride, unsigned channels)
{
    RpptDesc d{};
    d.numDims = channels > 1 ? 3 : 2;
    d.offsetInBytes = 0;
    d.dataType = RpptDataType::F32;
    d.n = B; d.c = channels; d.h = 1; d.w = stride / channels;
    d.strides.nStride = stride; d.strides.hStride = stride; d.strides.wStride = channels; d.strides.cStride = 1;
    d.layout = RpptLayout::NHW;
    return d;
}
RpptDesc desc_2d(int B, unsigned h, unsigned hStride, unsigned nStride, RpptLayout layout)
{
    RpptDesc d{};
    d.numDims = 3;
    d.dataType = RpptDataType::F32;
    d.n = B; d.c = 1; d.h = h; d.w = hStride;
    d.strides.nStride = nStride; d.strides.hStride = hStride; d.strides.wStride = 1; d.strides.cStride = 1;
    d.layout = layout;
    return d;
}
struct HostHandle {
    rppHandle_t h{};
    HostHandle(int B, int threads) { rppCreate(&h, B, threads, nullptr, RppBackend::RPP_HOST_BACKEND); }
    ~HostHandle() { rppDestroy(h, RppBackend::RPP_HOST_BACKEND); }
};
}  // namespace

extern "C" {

int ref_pre_emphasis(float *src, unsigned srcStride, float *dst, unsigned dstStride, int *len,
                     float *coeff, int B, int border, int threads)
{
    RpptDesc s = desc_1d(B, srcStride, 1), d = desc_1d(B, dstStride, 1);
    HostHandle hh(B, threads);
    return pre_emphasis_filter_host_tensor(src, &s, dst, &d, len, coeff, border, rpp::deref(hh.h));
}

int ref_mel_filter_bank(float *src, unsigned srcH, unsigned srcHStride, unsigned srcNStride, int *srcDims,
                        float *dst, unsigned dstHStride, unsigned dstNStride, int B,
                        float maxFreq, float minFreq, int formula, int numFilter, float sampleRate,
                        int normalize, int threads)
{
    RpptDesc s = desc_2d(B, srcH, srcHStride, srcNStride, RpptLayout::NFT);
    RpptDesc d = desc_2d(B, numFilter, dstHStride, dstNStride, RpptLayout::NFT);
    HostHandle hh(B, threads);
    return mel_filter_bank_host_tensor(src, &s, dst, &d, srcDims, maxFreq, minFreq,
                                       (RpptMelScaleFormula)formula, numFilter, sampleRate,
                                       normalize != 0, rpp::deref(hh.h));
}

// dims: RpptImagePatch {width, height} per clip
int ref_to_decibels(float *src, unsigned srcHStride, unsigned srcNStride, float *dst, unsigned dstHStride,
                    unsigned dstNStride, unsigned *dims, int B, float cutOffDB, float multiplier,
                    float referenceMagnitude, int threads)
{
    RpptDesc s = desc_2d(B, 1, srcHStride, srcNStride, RpptLayout::NFT);
    RpptDesc d = desc_2d(B, 1, dstHStride, dstNStride, RpptLayout::NFT);
    HostHandle hh(B, threads);
    return to_decibels_host_tensor(src, &s, dst, &d, reinterpret_cast<RpptImagePatch *>(dims), cutOffDB,
                                   multiplier, referenceMagnitude, rpp::deref(hh.h));
}

int ref_down_mixing(float *src, unsigned srcStride, unsigned channels, float *dst, unsigned dstStride,
                    int *srcDims, int B, int normalize, int threads)
{
    RpptDesc s = desc_1d(B, srcStride, channels), d = desc_1d(B, dstStride, 1);
    HostHandle hh(B, threads);
    return down_mixing_host_tensor(src, &s, dst, &d, srcDims, normalize != 0, rpp::deref(hh.h));
}

// lut: lookupSize floats; scale/center/lobes as RpptResamplingWindow (api/rppdefs.h:751-790)
int ref_resample(float *src, unsigned srcStride, float *dst, unsigned dstStride, float *inRate,
                 float *outRate, int *srcDims, int B, float scale, float center, int lobes, int coeffs,
                 int lookupSize, const float *lut, int threads)
{
    RpptDesc s = desc_1d(B, srcStride, 1), d = desc_1d(B, dstStride, 1);
    RpptResamplingWindow w;
    w.scale = scale; w.center = center; w.lobes = lobes; w.coeffs = coeffs; w.lookupSize = lookupSize;
    w.lookup.assign(lut, lut + lookupSize);
    w.pCenter = _mm_set1_ps(center);
    w.pScale = _mm_set1_ps(scale);
    HostHandle hh(B, threads);
    return resample_host_tensor(src, &s, dst, &d, inRate, outRate, srcDims, w, rpp::deref(hh.h));
}

int ref_non_silent_region_detection(float *src, unsigned srcStride, int *len, int B, float cutOffDB,
                                    int windowLength, float referencePower, int resetInterval, int *begin,
                                    int *length, int threads)
{
    RpptDesc s = desc_1d(B, srcStride, 1);
    HostHandle hh(B, threads);
    return non_silent_region_detection_host_tensor(src, &s, len, begin, length, cutOffDB, windowLength,
                                                   referencePower, resetInterval, rpp::deref(hh.h));
}

}  // extern "C"
