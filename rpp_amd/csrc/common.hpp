// common.hpp -- internal definitions shared by the gfx950 audio kernels and the C-ABI layer.
//
// Every translation unit is compiled with -ffp-contract=off: a*b+c is rounded twice unless
// the code asks for fmaf() explicitly.  This keeps the FP32 stages on the same rounding
// points as RPP's HOST kernels, which are built with -mno-fma (src/modules/CMakeLists.txt:105-124).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/rpp_audio.h"

#define RPPA_EXPORT extern "C" __attribute__((visibility("default")))

// HIP runtime failures become RPP_ERROR (the reference calls exit(-1), rppdefs.h:69-75).
#define RPPA_HIP(call)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            fprintf(stderr, "rpp_audio: %s failed: %s (%s:%d)\n", #call, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                                    \
            return RPP_ERROR;                                                               \
        }                                                                                   \
    } while (0)

namespace rppa {

constexpr int kWave = 64;

// Device-resident, lazily generated tables for one (nfft, windowLength, window) triple.
// basis[t * cols + c]: window-folded DFT columns (see tables.hip for the packing).
struct DftTables {
    int nfft = 0, W = 0, cols = 0;
    float *basis = nullptr;        // [W][cols] FP32
    // Folded split basis for the MFMA path (nfft a multiple of 4, even W, symmetric window).  With
    // the frame centre c = (W-1)/2, e_j = a[c+0.5+j] + a[c-0.5-j] and o_j = a[c+0.5+j] - a[c-0.5-j]
    // (j < W/2) give X_k = sum_j e_j w_j cos(2 pi k (j+0.5)/N) - i sum_j o_j w_j sin(...): half
    // the contraction length.  Only bins k < N/4 are stored: bin N/2 - k has
    // cos = (-1)^j sin_k and sin = (-1)^j cos_k at row j, so the same columns with the odd rows
    // negated give it; bin N/4 is formed separately from wHalf.  Columns: [0, N/4) w cos of bins
    // 0..N/4-1, [N/4, N/2) -w sin of the same bins.  Element (j, c) sits at
    // ((j / 8) * (N/2) + c) * 8 + j % 8 -- the operand order of a 32x32x16 MFMA fragment.
    uint16_t *foldHi = nullptr, *foldLo = nullptr;
    uint16_t *foldLo2 = nullptr;   // third term of the bf16 three-term split (splitKind 2), else null
    float *wHalf = nullptr;        // [W/2] symmetrised window half, w_j = w[W/2 + j]
    int Jp = 0;                    // W/2 padded to a multiple of 16 (zero rows)
    int splitKind = -1;            // format of foldHi/foldLo (-1 none, 0 bf16, 1 fp16 scaled by 4096,
                                   // 2 bf16 hi/mid/lo in foldHi/foldLo/foldLo2)
};

// Mel tables for one parameter set, built on the GPU (tables.hip).
struct MelTables {
    int numBins = 0, M = 0;
    int binStart = 0, binEnd = 0;
    float *weightsDown = nullptr;   // [numBins]
    int *intervals = nullptr;       // [numBins]  filter whose rising slope bin k is on, -1 unused
    float *normFactors = nullptr;   // [M]
    int *filterBins = nullptr;      // [M][2] first bin / one-past-last bin of filter m (gather form)
    float *dense = nullptr;         // [M][numBins] dense normalised weights (FP32)
    // bf16 hi/lo split of `dense` in MFMA operand order over power-plane positions p: element
    // (m, p) at ((p / 8) * Mpad + m) * 8 + p % 8.  Position p holds bin p for p <= Q and bin p - 3
    // for p >= Q + 4 (Q = (numBins - 1) / 2, i.e. N/4), positions Q+1..Q+3 hold nothing: the fused
    // kernel stores the upper half of the spectrum three places up so its 4-bin groups stay aligned.
    uint16_t *denseHi = nullptr, *denseLo = nullptr;
    int KP = 0;                     // numBins padded to a multiple of 32 (holds every position p above
                                    // when numBins - 1 is a multiple of 128)
    int Mpad = 0;                   // M padded to a multiple of 16
};

// Tables of the factored-DFT fused kernel (logmel_fft.hip) for one (nfft, W), built on the GPU in
// FP64: stage-1 and stage-2 MFMA operand fragments (fp16 hi/lo, scaled by 2^12), the window in
// polyphase order and the twiddles.  One allocation (`mem`).
struct FftTables {
    int nfft = 0, W = 0;
    uint16_t *a1 = nullptr, *a2 = nullptr;
    float *winP = nullptr, *tw = nullptr;
    float *tw2 = nullptr;   // nfft 1024: the radix-2 twiddles
    void *mem = nullptr;
};

// Mel weight fragments of the (16-filter tile, 32-bin chunk) pairs a filter bank touches, for the
// factored kernel: bf16 hi/lo in MFMA operand order, pairOff[tile * 16 + chunk] = fragment index or
// -1, pairOff[256] = nPairs.
// nfft 1024: each chunk has a run of pairs for bins 0 .. 255 and one for bins 512 - u (words
// pairOff[c] and pairOff[16 + c]), and w256 holds the weights of bin 256 (after pairOff's 257 ints).
struct MelFrag {
    uint16_t *frag = nullptr;
    int *pairOff = nullptr;
    int nPairs = 0;
    float *w256 = nullptr;
};

}  // namespace rppa

// The opaque handle of api/rpp.h.
struct rppHandle {
    RppBackend backend = RPP_HIP_BACKEND;
    int numThreads = 1;               // host threads of the rppt_*_host entry points
    int device = 0;
    hipStream_t stream = nullptr;
    size_t batchSize = 1;
    int numCU = 256;
    int maxLdsPerBlock = 64 * 1024;   // bytes of LDS one workgroup may allocate
    int maxLdsOptin = 64 * 1024;      // the same with the launch's explicit opt-in (dynamic LDS attribute)
    int numXCD = 8;

    // grow-only device workspace
    void *ws = nullptr;
    size_t wsBytes = 0;
    // grow-only buffer of the front end's two-launch form (the resampled clips and their lengths)
    void *feWs = nullptr;
    size_t feWsBytes = 0;

    // small pinned staging buffer for host-built metadata
    void *pinned = nullptr;
    size_t pinnedBytes = 0;

    std::map<std::tuple<int, int, uint64_t>, rppa::DftTables> dft;
    std::map<std::string, rppa::MelTables> mel;
    std::map<std::pair<int, int>, rppa::FftTables> fft;
    std::map<std::string, rppa::MelFrag> melFrag;

    RpptLogMelTiles tiles{0, 0};
};

namespace rppa {

// Returns a device pointer of at least `bytes` (grow-only); nullptr on failure or when the
// stream is being captured and the pool would have to grow.
void *workspace(rppHandle *h, size_t bytes);
// The same for the front end's resampled intermediate (a buffer of its own, so the log-mel launch
// that reads it can use the workspace).
void *front_workspace(rppHandle *h, size_t bytes);

// Window-folded DFT basis for (nfft, W, window); window == nullptr means RPP's Hann.
// splitKind: -1 FP32 only, 0 bf16 hi/lo, 1 fp16 hi/lo (scaled).
const DftTables *get_dft_tables(rppHandle *h, int nfft, int W, const float *window, int splitKind);

// Mel tables built the way the reference builds them (cpu/kernel/mel_filter_bank.cpp:55-112).
const MelTables *get_mel_tables(rppHandle *h, int numBins, float sampleRate, float maxFreq, float minFreq,
                                int formula, int M, int normalize);

// The same tables built into caller-provided device memory, enqueued on `st` with no allocation and
// no synchronisation.  DFT: `win` holds W floats of scratch (the window); t.basis ([W][cols] FP32)
// is filled when wantBasis, t.foldHi/foldLo (dft_fold_elems halves each) and t.wHalf (W/2 floats)
// when splitKind >= 0.  Mel: t.weightsDown/intervals [numBins], normFactors [M], filterBins [2M],
// dense [M][numBins], denseHi/denseLo (mel_dense_elems halves each), `range` 2 ints; binStart and
// binEnd stay unset (they are read back only by the cached builder).
size_t dft_fold_elems(int nfft, int W);
size_t mel_dense_elems(int numBins, int M);
hipError_t build_dft_tables_at(hipStream_t st, int nfft, int W, const float *window, int splitKind, bool wantBasis,
                               float *win, DftTables &t);
hipError_t build_mel_tables_at(hipStream_t st, int numBins, float sampleRate, float maxFreq, float minFreq, int formula,
                               int M, int normalize, int *range, MelTables &t);

// Factored-DFT fused kernel (logmel_fft.hip).
// The fused front end of the factored-DFT launch: interleaved clips at an integer multiple k of the
// output rate, down-mixed and resampled inside the staging (down_mix_resample_launch's arithmetic).
struct FrontEnd {
    const float *src;    // [B][stride] interleaved
    uint32_t stride;
    const int *dims;     // [2 b] samples per channel, [2 b + 1] channels
    float inRate, outRate;
    int k;               // inRate / outRate, 2 or 3
    int normalize;       // down-mix weights normalised
    const float *lut;    // the resampling window table
    float wScale, wCenter;
    int lookupSize;
};
// down-mix + resample of a batch at one rate pair (resample.hip's kernel, rates given once) into
// dst [B][dstStride], and each clip's resampled length into lengths[B]
hipError_t down_mix_resample_launch(hipStream_t st, const float *src, uint32_t srcStride, float *dst,
                                    uint32_t dstStride, unsigned B, const int *dims, float inRate, float outRate,
                                    const RpptResamplingWindow &w, int normalize, int *lengths);
struct FftLaunch {
    const RpptLogMelParams *p;
    const float *src;
    uint32_t srcStride;
    const long long *srcOff;
    const int *lengths;
    const int *begin;
    int B;
    float *dst;
    uint32_t dstNStride, hStride;
    const long long *dstOff;
    int ntf;
    uint32_t frameCap;
    int *clipMax;
    float minRatio, mult, invRef;
    const FftTables *ft;
    const MelFrag *mf;
    const FrontEnd *fe;  // null: the clips are given at the output rate
};
// the launch shape it supports (parameters only; the LDS check needs the mel pair count)
bool fft_eligible(const RpptLogMelParams *p, RpptLogMelPrecision precision, int B);
// the fused front end's resampling shapes: integer ratio k and window lobes
bool fft_front_eligible(int k, int lobes);
size_t fft_lds_bytes(int nfft, int H, int B, int nPairs, bool begin, bool packed, int *pitchOut, bool front = false,
                     int waves = 0);
// nfft 1024 runs only the common launch (power 2, NFT output, a fixed dB reference, a normal cut-off
// ratio) on clips given at the output rate
bool fft_r2_launchable(const RpptLogMelParams *p, int ntf, bool clipRef, bool frontEnd);
const FftTables *get_fft_tables(rppHandle *h, int nfft, int W);
const MelFrag *get_mel_frag(rppHandle *h, const MelTables *mt, const std::string &key);
// the same tables in caller-provided memory (no allocation): fft_table_bytes() bytes at `mem`
// (256-byte aligned) for W samples of window; mel_frag_bytes(M, nfft) for the fragments and
// mel_frag_word_bytes(nfft) for the chunk words (and nfft 1024's bin-256 weights after them)
size_t fft_table_bytes(int W, int nfft = 512);
void fft_tables_carve(char *mem, int W, FftTables &t, int nfft = 512);
hipError_t build_fft_tables_at(hipStream_t st, int W, FftTables &t);
size_t mel_frag_bytes(int M, int nfft = 512);
size_t mel_frag_word_bytes(int nfft = 512);
hipError_t build_mel_frag_at(hipStream_t st, const MelTables &mt, MelFrag &f);
RppStatus logmel_fft_launch(rppHandle *h, const FftLaunch &L);

inline unsigned ceil_div(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// RPP's frame count (cpu/kernel/spectrogram.cpp:47-52).
__host__ __device__ inline int num_windows(int length, int W, int H, bool center)
{
    if (!center) length -= W;
    return length / H + 1;
}

// RPP's reflect index over [0, L) (cpu/kernel/spectrogram.cpp:55-69): mirror about 0 and L-1
// without repeating the edge, iterated.  That map is periodic with period 2(L-1), so it is
// evaluated in closed form here; L < 2 maps everything to L-1 as the reference does.
__host__ __device__ inline int idx_reflect(int loc, int L)
{
    if (L < 2) return L - 1;
    const int P = 2 * (L - 1);
    int m = loc % P;
    if (m < 0) m += P;
    return m < L ? m : P - m;
}

}  // namespace rppa
