// A C++ caller of librpp_audio.so written the way a user of RPP's audio GPU path writes one: the
// example of INTEGRATION.md, compiled against include/rpp_audio.h and linked against the library by
// tests/test_boundary.py.  The static_asserts pin the descriptor layouts to the reference's
// (api/rppdefs.h:570-589, :645-649, :751-790 on x86-64) and the enum values a binary caller passes.
// main() resolves every entry point and runs the HOST-backend chain on the CPU (run_host_chain);
// the GPU sequence (run_chain) is only linked, not called.
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>

#include "rpp_audio.h"

static_assert(sizeof(RpptDesc) == 56, "RpptDesc layout");
static_assert(offsetof(RpptDesc, numDims) == 0 && offsetof(RpptDesc, offsetInBytes) == 8, "RpptDesc layout");
static_assert(offsetof(RpptDesc, dataType) == 12 && offsetof(RpptDesc, n) == 16, "RpptDesc layout");
static_assert(offsetof(RpptDesc, c) == 20 && offsetof(RpptDesc, h) == 24 && offsetof(RpptDesc, w) == 28,
              "RpptDesc layout");
static_assert(offsetof(RpptDesc, strides) == 32 && offsetof(RpptDesc, layout) == 48, "RpptDesc layout");
static_assert(sizeof(RpptStrides) == 16, "RpptStrides layout");
static_assert(sizeof(RpptImagePatch) == 8 && offsetof(RpptImagePatch, height) == 4, "RpptImagePatch layout");
static_assert(sizeof(RpptResamplingWindow) == 96, "RpptResamplingWindow layout");
static_assert(offsetof(RpptResamplingWindow, scale) == 0 && offsetof(RpptResamplingWindow, center) == 4,
              "RpptResamplingWindow layout");
static_assert(offsetof(RpptResamplingWindow, lobes) == 8 && offsetof(RpptResamplingWindow, coeffs) == 12,
              "RpptResamplingWindow layout");
static_assert(offsetof(RpptResamplingWindow, lookupSize) == 16 && offsetof(RpptResamplingWindow, lookupPinned) == 24,
              "RpptResamplingWindow layout");
static_assert(offsetof(RpptResamplingWindow, lookup) == 32 && offsetof(RpptResamplingWindow, pCenter) == 64 &&
                  offsetof(RpptResamplingWindow, pScale) == 80,
              "RpptResamplingWindow layout");
static_assert(sizeof(RpptLayout) == 4 && NFT == 5 && NTF == 6, "RpptLayout values");
static_assert(F32 == 1 && ZERO == 0 && CLAMP == 1 && REFLECT == 2 && SLANEY == 0 && HTK == 1, "enum values");
static_assert(RPP_HOST_BACKEND == 0 && RPP_HIP_BACKEND == 1, "RppBackend values");
static_assert(RPP_ERROR_NOT_ENOUGH_MEMORY == -16 && RPP_ERROR_INVALID_DST_DIMS == -24, "status codes");

// The INTEGRATION.md call sequence.  Buffers are the caller's: device tensors, per-clip arrays in
// device or pinned memory.
int run_chain(void *stream, int batch, float *d_src, RpptDesc &srcDesc, float *d_pre, float *d_spec,
              RpptDesc &specDesc, float *d_mel, RpptDesc &melDesc, float *d_db, RpptDesc &dbDesc, int *d_len,
              float *d_coeff, int *d_melDims, RpptImagePatch *d_dbDims, float *d_stereo, RpptDesc &stereoDesc,
              float *d_mono16k, RpptDesc &monoDesc, float *d_inRate, float *d_outRate, int *d_dims,
              RpptResamplingWindow &window, int *d_begin, int *d_regionLen)
{
    rppHandle_t h;
    if (rppCreate(&h, batch, 0, stream, RPP_HIP_BACKEND) != rppStatusSuccess) return 1;
    int bad = 0;
    bad |= rppt_pre_emphasis_filter_gpu(d_src, &srcDesc, d_pre, &srcDesc, d_len, d_coeff, CLAMP, h);
    bad |= rppt_spectrogram_gpu(d_pre, &srcDesc, d_spec, &specDesc, d_len, true, true, nullptr, 512, 2, 400, 160, h);
    bad |= rppt_mel_filter_bank_gpu(d_spec, &specDesc, d_mel, &melDesc, d_melDims, 8000.f, 0.f, SLANEY, 80, 16000.f,
                                    true, h);
    bad |= rppt_to_decibels_gpu(d_mel, &melDesc, d_db, &melDesc, d_dbDims, -80.f, 10.f, 1.f, h);
    RpptLogMelParams p = {0.97f, CLAMP, 512, 400, 160, 1, 1, 2, 16000.f, 0.f, 8000.f, 80, SLANEY, 1, -80.f, 10.f, 1.f};
    bad |= rppt_logmel_fused_gpu(d_src, &srcDesc, d_db, &dbDesc, d_len, &p, RPPT_LOGMEL_FP16X3, h);
    bad |= rppt_down_mix_resample_gpu(d_stereo, &stereoDesc, d_mono16k, &monoDesc, d_inRate, d_outRate, d_dims, false,
                                      window, h);
    bad |= rppt_non_silent_region_detection_gpu(d_src, &srcDesc, d_len, d_begin, d_regionLen, -60.f, 2048, 0.f, 8192, h);
    bad |= rppt_logmel_fused_region_gpu(d_src, &srcDesc, d_db, &dbDesc, d_begin, d_regionLen, &p, RPPT_LOGMEL_FP16X3,
                                        h);
    rppDestroy(h, RPP_HIP_BACKEND);
    return bad != 0;
}

// The HOST backend on host buffers: the same chain through rppt_*_host, run on the CPU.  One clip
// of a 1 kHz tone at 16 kHz; returns 0 when every call succeeds and the dB peak (reference = the
// clip's maximum) is 0 dB in the filter holding 1 kHz.
int run_host_chain()
{
    const int L = 16000, F = L / 160 + 1, K = 257, M = 80;
    static float x[L], pre[L], spec[K * F], mel[M * F], db[M * F];
    for (int t = 0; t < L; t++) x[t] = 0.5f * std::sin(2.0f * 3.14159265f * 1000.0f * t / 16000.0f);
    int len = L, melDims[2] = {K, F};
    float coeff = 0.97f;
    RpptImagePatch dbDims = {static_cast<Rpp32u>(F), static_cast<Rpp32u>(M)};
    RpptDesc sd = {2, 0, F32, 1, 1, 1, static_cast<Rpp32u>(L), {static_cast<Rpp32u>(L), static_cast<Rpp32u>(L), 1, 1}, NHW};
    RpptDesc specD = {3, 0, F32, 1, 1, K, static_cast<Rpp32u>(F), {static_cast<Rpp32u>(K * F), static_cast<Rpp32u>(K * F), static_cast<Rpp32u>(F), 1}, NFT};
    RpptDesc melD = {3, 0, F32, 1, 1, M, static_cast<Rpp32u>(F), {static_cast<Rpp32u>(M * F), static_cast<Rpp32u>(M * F), static_cast<Rpp32u>(F), 1}, NFT};
    rppHandle_t h;
    if (rppCreate(&h, 1, 2, nullptr, RPP_HOST_BACKEND) != rppStatusSuccess) return 1;
    int bad = 0;
    bad |= rppt_pre_emphasis_filter_host(x, &sd, pre, &sd, &len, &coeff, CLAMP, h);
    bad |= rppt_spectrogram_host(pre, &sd, spec, &specD, &len, true, true, nullptr, 512, 2, 400, 160, h);
    bad |= rppt_mel_filter_bank_host(spec, &specD, mel, &melD, melDims, 8000.f, 0.f, SLANEY, M, 16000.f, true, h);
    bad |= rppt_to_decibels_host(mel, &melD, db, &melD, &dbDims, -80.f, 10.f, 0.f, h);
    rppDestroy(h, RPP_HOST_BACKEND);
    if (bad) return 1;
    int best = 0;
    for (int m = 0; m < M; m++)
        if (db[m * F + F / 2] > db[best * F + F / 2]) best = m;
    std::printf("host chain: loudest filter %d at %.3f dB\n", best, db[best * F + F / 2]);
    return db[best * F + F / 2] > -1.0f && best > 10 && best < 40 ? 0 : 1;
}

int main(int argc, char **argv)
{
    (void)argv;
    // every entry point's address is taken, so the link resolves all of them
    const void *syms[] = {
        (const void *)&rppCreate, (const void *)&rppDestroy, (const void *)&rppSetStream,
        (const void *)&rppGetStream, (const void *)&rppSetBatchSize, (const void *)&rppGetBatchSize,
        (const void *)&rppt_pre_emphasis_filter_gpu, (const void *)&rppt_spectrogram_gpu,
        (const void *)&rppt_mel_filter_bank_gpu, (const void *)&rppt_to_decibels_gpu,
        (const void *)&rppt_down_mixing_gpu, (const void *)&rppt_resample_gpu,
        (const void *)&rppt_non_silent_region_detection_gpu, (const void *)&rppt_down_mix_resample_gpu,
        (const void *)&rppt_logmel_fused_gpu, (const void *)&rppt_logmel_fused_region_gpu,
        (const void *)&rppt_logmel_fused_packed_gpu, (const void *)&rppt_logmel_chain_gpu,
        (const void *)&rppt_logmel_set_tiles, (const void *)&rppt_logmel_get_tiles,
        (const void *)&rppt_logmel_autotune, (const void *)&rppa_server_create, (const void *)&rppa_server_destroy,
        (const void *)&rppa_server_submit, (const void *)&rppa_server_wait, (const void *)&rppa_server_drain,
        (const void *)&rppa_server_stats, (const void *)&run_chain,
        (const void *)&rppt_pre_emphasis_filter_host, (const void *)&rppt_spectrogram_host,
        (const void *)&rppt_mel_filter_bank_host, (const void *)&rppt_to_decibels_host,
        (const void *)&rppt_down_mixing_host, (const void *)&rppt_resample_host,
        (const void *)&rppt_non_silent_region_detection_host, (const void *)&run_host_chain};
    int n = 0;
    for (const void *s : syms) n += s != nullptr;
    std::printf("%d entry points resolved\n", n);
    if (run_host_chain() != 0) return 3;
    return argc > 1 ? 2 : 0;
}
