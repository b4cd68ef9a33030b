// tables.hip -- window, window-folded DFT basis and mel filter tables, generated on the GPU in FP64.
//
// DFT basis column packing (nfft = N, K = N/2 + 1 bins, W window samples):
//   column 0        w[t] * cos(0)              -> bin 0 (DC, real only)
//   column 1        w[t] * cos(pi t)           -> bin N/2 (Nyquist, real only; zero when N is odd)
//   columns 2k,2k+1 w[t] * cos(2 pi k t / N), -w[t] * sin(2 pi k t / N)   for 1 <= k <= (N-1)/2
// so cols = N (+1 when N is odd) and power[k] = C[2k]^2 + C[2k+1]^2, power[0] = C[0]^2,
// power[N/2] = C[1]^2.  The reference places the frame at offset (N-W)/2 of the FFT buffer
// (cpu/kernel/spectrogram.cpp:197-212); that only rotates phase, so the basis uses origin 0.
#include "common.hpp"

#include <cstring>

namespace rppa {
namespace {

constexpr double kPi = 3.14159265358979323846;

// RPP's Hann with a half-sample shift (cpu/kernel/spectrogram.cpp:36-44), FP64 -> FP32.
__global__ void hann_kernel(int W, float *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= W) return;
    double a = (2.0 * kPi) / W;
    out[t] = static_cast<float>(0.5 * (1.0 - cos(a * (t + 0.5))));
}

__device__ inline uint16_t f32_to_bf16_rne(float v)
{
    uint32_t u = __float_as_uint(v);
    uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(r >> 16);
}
__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ inline uint16_t f32_to_f16_rne(float v)
{
    _Float16 h = static_cast<_Float16>(v);
    uint16_t r;
    __builtin_memcpy(&r, &h, 2);
    return r;
}
__device__ inline float f16_to_f32(uint16_t u)
{
    _Float16 h;
    __builtin_memcpy(&h, &u, 2);
    return static_cast<float>(h);
}

// One thread per (t, c): basis[t][c] = window-folded DFT column c at sample t, FP64 -> FP32.
__global__ void dft_basis_kernel(int N, int W, int cols, const float *__restrict__ win, float *__restrict__ basis)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    int t = blockIdx.y;
    if (c >= cols || t >= W) return;
    double w = static_cast<double>(win[t]);
    int k;
    bool isSin = false;
    if (c == 0) k = 0;
    else if (c == 1) k = (N % 2 == 0) ? N / 2 : -1;
    else { k = c >> 1; isSin = (c & 1) != 0; }
    double v = 0.0;
    if (k >= 0) {
        long long kt = (static_cast<long long>(k) * t) % N;
        double th = 2.0 * kPi * static_cast<double>(kt) / N;
        v = isSin ? -w * sin(th) : w * cos(th);
    }
    basis[static_cast<size_t>(t) * cols + c] = static_cast<float>(v);
}

// One thread per (j, c) of the folded split basis (see common.hpp).  The window is symmetrised
// (mean of w[W/2+j] and w[W/2-1-j], equal for RPP's Hann up to FP32 rounding).  Angles are
// reduced exactly: k (2j+1) mod 2N over 2N.
__global__ void fold_basis_kernel(int N, int W, int Jp, const float *__restrict__ win, uint16_t *__restrict__ hi,
                                  uint16_t *__restrict__ lo, uint16_t *__restrict__ lo2, float *__restrict__ wHalf,
                                  int splitKind, float splitScale)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int NC = N / 2, Q = N / 4;
    if (c >= NC || j >= Jp) return;
    const int half = W / 2;
    double v = 0.0;
    if (j < half) {
        const double w = 0.5 * (static_cast<double>(win[half + j]) + static_cast<double>(win[half - 1 - j]));
        if (c == 0) wHalf[j] = static_cast<float>(w);
        const bool isSin = c >= Q;
        const long long k = isSin ? c - Q : c;
        const long long m = (k * (2LL * j + 1)) % (2LL * N);
        const double th = kPi * static_cast<double>(m) / N;
        v = isSin ? -w * sin(th) : w * cos(th);
    }
    const double s = v * splitScale;
    uint16_t h, l, l2 = 0;
    if (splitKind == 2) {
        // three bf16 terms, each the rounded remainder of the ones before (remainders in FP64)
        h = f32_to_bf16_rne(static_cast<float>(s));
        const double r1 = s - bf16_to_f32(h);
        l = f32_to_bf16_rne(static_cast<float>(r1));
        l2 = f32_to_bf16_rne(static_cast<float>(r1 - bf16_to_f32(l)));
    } else if (splitKind == 0) {
        h = f32_to_bf16_rne(static_cast<float>(s));
        l = f32_to_bf16_rne(static_cast<float>(s - bf16_to_f32(h)));
    } else {
        h = f32_to_f16_rne(static_cast<float>(s));
        l = f32_to_f16_rne(static_cast<float>(s - f16_to_f32(h)));
    }
    const size_t o = (static_cast<size_t>(j / 8) * NC + c) * 8 + (j % 8);
    hi[o] = h;
    lo[o] = l;
    if (splitKind == 2) lo2[o] = l2;
}

// bf16 hi/lo split of the dense mel matrix in MFMA operand order (see common.hpp).
__global__ void mel_split_kernel(const float *__restrict__ dense, int M, int numBins, int Mpad, int KP,
                                 uint16_t *__restrict__ hi, uint16_t *__restrict__ lo)
{
    int m = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;   // power-plane position
    if (m >= Mpad || k >= KP) return;
    const int Q = (numBins - 1) / 2;
    const int bin = k <= Q ? k : (k >= Q + 4 ? k - 3 : -1);
    float v = (m < M && bin >= 0 && bin < numBins) ? dense[static_cast<size_t>(m) * numBins + bin] : 0.0f;
    uint16_t h = f32_to_bf16_rne(v);
    uint16_t l = f32_to_bf16_rne(v - bf16_to_f32(h));
    size_t o = (static_cast<size_t>(k / 8) * Mpad + m) * 8 + (k % 8);
    hi[o] = h;
    lo[o] = l;
}

// ---------------------------------------------------------------------------------------------
// Mel scales as RPP defines them: FP32 member functions (api/rppdefs.h:795-851).  logf/expf are
// evaluated as FP64 log/exp rounded once to FP32 (the correctly rounded FP32 result).
// ---------------------------------------------------------------------------------------------
__device__ inline float logf_cr(float x) { return static_cast<float>(log(static_cast<double>(x))); }
__device__ inline float expf_cr(float x) { return static_cast<float>(exp(static_cast<double>(x))); }

struct MelScale {
    int formula;  // 0 Slaney, 1 HTK
    __device__ float minLogMel() const { return (1000.0f - 0.0f) / 66.666667f; }
    __device__ float hz_to_mel(float hz) const
    {
        if (formula == 1) return 1127.0f * logf_cr(1.0f + (hz / 700.0f));
        const float invStepLog = 1.0f / 0.068751777f, invFsp = 1.0f / 66.666667f;
        if (hz >= 1000.0f) return minLogMel() + logf_cr(hz * 0.001f) * invStepLog;
        return (hz - 0.0f) * invFsp;
    }
    __device__ float mel_to_hz(float mel) const
    {
        if (formula == 1) return 700.0f * (expf_cr(mel / 1127.0f) - 1.0f);
        if (mel >= minLogMel()) return 1000.0f * expf_cr(0.068751777f * (mel - minLogMel()));
        return 0.0f + mel * 66.666667f;
    }
};

// Table build.  The mel points and per-bin weights follow the reference's loop exactly on one
// thread (cpu/kernel/mel_filter_bank.cpp:55-112): FP64 mel points accumulated by repeated
// addition, FP32 scale functions, FP64 slopes rounded to FP32 weights.  The initialisation, the
// per-filter bin ranges and the dense matrix are independent per element and run over the block.
__global__ void __launch_bounds__(256)
mel_tables_kernel(int numBins, float sampleRate, float maxFreqVal, float minFreq, int formula, int M, int normalize,
                  float *weightsDown, int *intervals, float *normFactors, int *filterBins, float *dense,
                  int *binRange)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    __shared__ int range[2];
    for (int m = tid; m < M; m += nt) {
        normFactors[m] = 1.0f;
        filterBins[2 * m] = 0x7fffffff;   // first bin (min), then last bin + 1 (max)
        filterBins[2 * m + 1] = 0;
    }
    for (int k = tid; k < numBins; k += nt) { weightsDown[k] = 0.0f; intervals[k] = -1; }
    for (int i = tid; i < M * numBins; i += nt) dense[i] = 0.0f;
    __syncthreads();
    if (tid == 0) {
        MelScale ms{formula};
        float maxFreq = (maxFreqVal == 0) ? sampleRate / 2 : maxFreqVal;
        double melLow = ms.hz_to_mel(minFreq);
        double melHigh = ms.hz_to_mel(maxFreq);
        double melStep = (melHigh - melLow) / (M + 1);
        int nfft = (numBins - 1) * 2;
        double hzStep = static_cast<double>(sampleRate) / nfft;
        double invHzStep = 1.0 / hzStep;
        int binStart = static_cast<int>(ceil(minFreq * invHzStep));
        int binEnd = static_cast<int>(ceil(maxFreq * invHzStep));
        if (binEnd > numBins) binEnd = numBins;
        int bin = binStart;
        double mel0 = melLow, mel1 = melLow + melStep;
        double fIter = bin * hzStep;
        for (int iv = 0; iv < M + 1; iv++, mel0 = mel1, mel1 += melStep) {
            double f0 = ms.mel_to_hz(static_cast<float>(mel0));
            double f1 = ms.mel_to_hz(static_cast<float>(iv == M ? melHigh : mel1));
            double slope = 1. / (f1 - f0);
            if (normalize && iv < M) {
                double f2 = ms.mel_to_hz(static_cast<float>(mel1 + melStep));
                normFactors[iv] = static_cast<float>(2.0 / (f2 - f0));
            }
            for (; bin < binEnd && fIter < f1; bin++, fIter = bin * hzStep) {
                weightsDown[bin] = static_cast<float>((f1 - fIter) * slope);
                intervals[bin] = iv;
            }
        }
        binRange[0] = range[0] = binStart;
        binRange[1] = range[1] = binEnd;
    }
    __syncthreads();
    // bin k feeds filter `up` with weight 1 - weightsDown[k] and filter up - 1 with weightsDown[k];
    // each filter's bin range is [first, last + 1) over the bins that feed it, [0, 0) if none
    for (int k = range[0] + tid; k < range[1]; k += nt) {
        const int up = intervals[k], down = up - 1;
        const float wDown = weightsDown[k], wUp = 1.0f - weightsDown[k];
        if (down >= 0 && down < M) {
            dense[down * numBins + k] = normalize ? wDown * normFactors[down] : wDown;
            atomicMin(&filterBins[2 * down], k);
            atomicMax(&filterBins[2 * down + 1], k + 1);
        }
        if (up >= 0 && up < M) {
            dense[up * numBins + k] = normalize ? wUp * normFactors[up] : wUp;
            atomicMin(&filterBins[2 * up], k);
            atomicMax(&filterBins[2 * up + 1], k + 1);
        }
    }
    __syncthreads();
    for (int m = tid; m < M; m += nt)
        if (filterBins[2 * m] == 0x7fffffff) filterBins[2 * m] = 0;
}

uint64_t fnv1a(const void *p, size_t n)
{
    const unsigned char *b = static_cast<const unsigned char *>(p);
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

}  // namespace

size_t dft_fold_elems(int nfft, int W) { return static_cast<size_t>((W / 2 + 15) / 16 * 16) * (nfft / 2); }

size_t mel_dense_elems(int numBins, int M)
{
    return static_cast<size_t>((numBins + 31) / 32 * 32) * ((M + 15) / 16 * 16);
}

hipError_t build_dft_tables_at(hipStream_t st, int nfft, int W, const float *window, int splitKind, bool wantBasis,
                               float *win, DftTables &t)
{
    t.nfft = nfft;
    t.W = W;
    t.cols = nfft + (nfft & 1);
    if (window) {
        const hipError_t e = hipMemcpyAsync(win, window, sizeof(float) * W, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
    } else {
        hipLaunchKernelGGL(hann_kernel, dim3(ceil_div(W, 256)), dim3(256), 0, st, W, win);
    }
    if (wantBasis)
        hipLaunchKernelGGL(dft_basis_kernel, dim3(ceil_div(t.cols, 256), W), dim3(256), 0, st, nfft, W, t.cols, win,
                           t.basis);
    if (splitKind >= 0) {
        t.Jp = (W / 2 + 15) / 16 * 16;
        t.splitKind = splitKind;
        const float splitScale = (splitKind == 1) ? 4096.0f : 1.0f;
        hipLaunchKernelGGL(fold_basis_kernel, dim3(ceil_div(nfft / 2, 256), t.Jp), dim3(256), 0, st, nfft, W, t.Jp,
                           win, t.foldHi, t.foldLo, t.foldLo2, t.wHalf, splitKind, splitScale);
    }
    return hipGetLastError();
}

hipError_t build_mel_tables_at(hipStream_t st, int numBins, float sampleRate, float maxFreq, float minFreq, int formula,
                               int M, int normalize, int *range, MelTables &t)
{
    t.numBins = numBins;
    t.M = M;
    t.KP = (numBins + 31) / 32 * 32;
    t.Mpad = (M + 15) / 16 * 16;
    hipLaunchKernelGGL(mel_tables_kernel, dim3(1), dim3(256), 0, st, numBins, sampleRate, maxFreq, minFreq, formula, M,
                       normalize, t.weightsDown, t.intervals, t.normFactors, t.filterBins, t.dense, range);
    hipLaunchKernelGGL(mel_split_kernel, dim3(ceil_div(t.Mpad, 64), t.KP), dim3(64), 0, st, t.dense, M, numBins,
                       t.Mpad, t.KP, t.denseHi, t.denseLo);
    return hipGetLastError();
}

const DftTables *get_dft_tables(rppHandle *h, int nfft, int W, const float *window, int splitKind)
{
    uint64_t key = window ? fnv1a(window, sizeof(float) * W) : 0;
    auto k = std::make_tuple(nfft, W, key);
    auto it = h->dft.find(k);
    if (it != h->dft.end() && (splitKind < 0 || it->second.splitKind == splitKind)) return &it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return nullptr;
    if (splitKind >= 0 && ((nfft & 1) || (W & 1))) return nullptr;  // the fold needs even nfft and W

    DftTables t;
    const bool fresh = it == h->dft.end();
    if (!fresh) t = it->second;
    float *win = nullptr;
    if (hipMalloc(&win, sizeof(float) * W) != hipSuccess) return nullptr;
    if (fresh && hipMalloc(&t.basis, sizeof(float) * static_cast<size_t>(W) * (nfft + (nfft & 1))) != hipSuccess)
        return nullptr;
    if (splitKind >= 0) {
        if (t.foldHi) { (void)hipFree(t.foldHi); (void)hipFree(t.foldLo); (void)hipFree(t.wHalf); }
        if (t.foldLo2) (void)hipFree(t.foldLo2);
        t.foldLo2 = nullptr;
        const size_t n = dft_fold_elems(nfft, W);
        if (hipMalloc(&t.foldHi, n * 2) != hipSuccess || hipMalloc(&t.foldLo, n * 2) != hipSuccess ||
            hipMalloc(&t.wHalf, sizeof(float) * (W / 2)) != hipSuccess ||
            (splitKind == 2 && hipMalloc(&t.foldLo2, n * 2) != hipSuccess))
            return nullptr;
    }
    if (build_dft_tables_at(h->stream, nfft, W, window, splitKind, fresh, win, t) != hipSuccess) return nullptr;
    if (hipStreamSynchronize(h->stream) != hipSuccess) return nullptr;
    (void)hipFree(win);
    h->dft[k] = t;
    return &h->dft[k];
}

const MelTables *get_mel_tables(rppHandle *h, int numBins, float sampleRate, float maxFreq, float minFreq,
                                int formula, int M, int normalize)
{
    char keyb[160];
    snprintf(keyb, sizeof(keyb), "%d/%a/%a/%a/%d/%d/%d", numBins, sampleRate, maxFreq, minFreq, formula, M, normalize);
    std::string key(keyb);
    auto it = h->mel.find(key);
    if (it != h->mel.end()) return &it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return nullptr;

    MelTables t;
    int *range = nullptr;
    const size_t ns = mel_dense_elems(numBins, M);
    if (hipMalloc(&t.weightsDown, sizeof(float) * numBins) != hipSuccess ||
        hipMalloc(&t.intervals, sizeof(int) * numBins) != hipSuccess ||
        hipMalloc(&t.normFactors, sizeof(float) * M) != hipSuccess ||
        hipMalloc(&t.filterBins, sizeof(int) * 2 * M) != hipSuccess ||
        hipMalloc(&t.dense, sizeof(float) * static_cast<size_t>(M) * numBins) != hipSuccess ||
        hipMalloc(&t.denseHi, ns * 2) != hipSuccess || hipMalloc(&t.denseLo, ns * 2) != hipSuccess ||
        hipMalloc(&range, sizeof(int) * 2) != hipSuccess)
        return nullptr;
    if (build_mel_tables_at(h->stream, numBins, sampleRate, maxFreq, minFreq, formula, M, normalize, range, t) != hipSuccess)
        return nullptr;
    int hr[2] = {0, 0};
    if (hipMemcpyAsync(hr, range, sizeof(hr), hipMemcpyDeviceToHost, h->stream) != hipSuccess) return nullptr;
    if (hipStreamSynchronize(h->stream) != hipSuccess) return nullptr;
    (void)hipFree(range);
    t.binStart = hr[0];
    t.binEnd = hr[1];
    h->mel[key] = t;
    return &h->mel[key];
}

}  // namespace rppa

// Debug/test export: copies the mel tables of a parameter set to host arrays.
RPPA_EXPORT RppStatus rppa_debug_mel_tables(rppHandle_t handle, int numBins, float sampleRate, float maxFreq,
                                            float minFreq, int formula, int M, int normalize, float *weightsDown,
                                            int *intervals, float *normFactors, int *binRange)
{
    if (!handle) return RPP_ERROR_INVALID_ARGUMENTS;
    const rppa::MelTables *t = rppa::get_mel_tables(handle, numBins, sampleRate, maxFreq, minFreq, formula, M, normalize);
    if (!t) return RPP_ERROR;
    RPPA_HIP(hipMemcpy(weightsDown, t->weightsDown, sizeof(float) * numBins, hipMemcpyDeviceToHost));
    RPPA_HIP(hipMemcpy(intervals, t->intervals, sizeof(int) * numBins, hipMemcpyDeviceToHost));
    RPPA_HIP(hipMemcpy(normFactors, t->normFactors, sizeof(float) * M, hipMemcpyDeviceToHost));
    binRange[0] = t->binStart;
    binRange[1] = t->binEnd;
    return RPP_SUCCESS;
}
