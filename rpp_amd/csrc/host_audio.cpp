// host_audio.cpp -- the HOST backend: rppt_*_host entry points of
// api/rppt_tensor_audio_augmentations.h on the CPU cores, for callers holding host buffers with a
// handle made by rppCreate(..., RPP_HOST_BACKEND) (RPP's default backend, api/rpp.h:132).
//
// Semantics are those of RPP's HOST kernels (src/modules/tensor/cpu/kernel/*.cpp), including their
// rounding points where the outputs are exact functions of the inputs: pre-emphasis, mel filter
// bank, to_decibels, down-mixing, resampling and region detection give RPP HOST's bits (products
// rounded before the add: every file here is built with -ffp-contract=off, as RPP builds these
// kernels with -mno-fma).  The spectrogram's FFT is this file's own split real FFT (RPP links the
// prebuilt FFTS library), so it agrees with RPP HOST to FP32 rounding, not bit for bit.
//
// Work is split over clips as RPP does (one OpenMP iteration per clip): here a pool of
// std::threads, as many as the handle's thread count, takes clips from a shared counter.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

#include "common.hpp"

namespace rppa {
namespace host {
namespace {

// Runs fn(b) for b in [0, n) on up to `threads` threads (the calling thread is one of them).
template <class Fn>
void for_each_clip(int n, int threads, Fn &&fn)
{
    threads = std::max(1, std::min(threads, n));
    if (threads == 1) {
        for (int b = 0; b < n; b++) fn(b);
        return;
    }
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int b = next.fetch_add(1); b < n; b = next.fetch_add(1)) fn(b);
    };
    std::vector<std::thread> pool;
    pool.reserve(threads - 1);
    for (int t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
}

int threads_of(rppHandle_t h) { return h->numThreads > 0 ? h->numThreads : 1; }

// ------------------------------------------------------------------ real FFT
// n = 2^p real samples -> bins 0..n/2, as a complex FFT of the m = n/2 even/odd pairs followed by
// the split step X[k] = E[k] + e^{-2 pi i k/n} O[k], with E and O recovered from Z[k] and
// conj(Z[m-k]).  The complex FFT is iterative radix-2 decimation in time; stage s (butterfly span
// 2^s) reads its 2^(s-1) twiddles from a contiguous run of `tw`.
class RealFft {
public:
    explicit RealFft(int n) : n_(n), m_(n / 2)
    {
        int lg = 0;
        while ((1 << lg) < m_) lg++;
        rev_.resize(m_);
        for (int i = 0; i < m_; i++) {
            int r = 0;
            for (int b = 0; b < lg; b++)
                if (i >> b & 1) r |= 1 << (lg - 1 - b);
            rev_[i] = r;
        }
        tw_.resize(std::max(m_, 1));
        for (int half = 1; half < m_; half <<= 1)
            for (int j = 0; j < half; j++) {
                const double a = -M_PI * j / half;
                tw_[half + j] = std::complex<float>(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
            }
        split_.resize(m_ + 1);
        for (int k = 0; k <= m_; k++) {
            const double a = -2.0 * M_PI * k / n;
            split_[k] = std::complex<float>(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
        }
    }

    // in: n reals; out: n/2 + 1 bins; z: m complex of scratch
    void run(const float *in, std::complex<float> *out, std::complex<float> *z) const
    {
        for (int i = 0; i < m_; i++) z[rev_[i]] = std::complex<float>(in[2 * i], in[2 * i + 1]);
        for (int half = 1; half < m_; half <<= 1) {
            const std::complex<float> *w = &tw_[half];
            for (int s = 0; s < m_; s += 2 * half) {
                std::complex<float> *lo = z + s, *hi = z + s + half;
                for (int j = 0; j < half; j++) {
                    const float tr = hi[j].real() * w[j].real() - hi[j].imag() * w[j].imag();
                    const float ti = hi[j].real() * w[j].imag() + hi[j].imag() * w[j].real();
                    const float ar = lo[j].real(), ai = lo[j].imag();
                    hi[j] = std::complex<float>(ar - tr, ai - ti);
                    lo[j] = std::complex<float>(ar + tr, ai + ti);
                }
            }
        }
        out[0] = std::complex<float>(z[0].real() + z[0].imag(), 0.0f);
        out[m_] = std::complex<float>(z[0].real() - z[0].imag(), 0.0f);
        for (int k = 1; k < m_; k++) {
            const std::complex<float> a = z[k], c = std::conj(z[m_ - k]);
            const float er = 0.5f * (a.real() + c.real()), ei = 0.5f * (a.imag() + c.imag());
            const float orr = 0.5f * (a.imag() - c.imag()), oi = -0.5f * (a.real() - c.real());
            const std::complex<float> w = split_[k];
            out[k] = std::complex<float>(er + (orr * w.real() - oi * w.imag()), ei + (orr * w.imag() + oi * w.real()));
        }
    }

private:
    int n_, m_;
    std::vector<int> rev_;
    std::vector<std::complex<float>> tw_, split_;
};

// Direct DFT for lengths the split FFT does not take (not a power of two, or below 4): bins
// 0..n/2 of n samples, twiddles from an exact table of e^{-2 pi i j/n}, sums in double.
class DirectDft {
public:
    explicit DirectDft(int n) : n_(n), c_(n), s_(n)
    {
        for (int j = 0; j < n; j++) {
            c_[j] = std::cos(2.0 * M_PI * j / n);
            s_[j] = std::sin(2.0 * M_PI * j / n);
        }
    }
    void run(const float *in, std::complex<float> *out) const
    {
        for (int k = 0; k <= n_ / 2; k++) {
            double re = 0.0, im = 0.0;
            long j = 0;
            for (int t = 0; t < n_; t++, j += k) {
                if (j >= n_) j -= n_;
                re += in[t] * c_[j];
                im -= in[t] * s_[j];
            }
            out[k] = std::complex<float>(static_cast<float>(re), static_cast<float>(im));
        }
    }

private:
    int n_;
    std::vector<double> c_, s_;
};

int frames_of(int L, int W, int H, bool center) { return num_windows(L, W, H, center); }

// ------------------------------------------------------------------ mel scales (api/rppdefs.h:795-851)
// FP32 member functions as RPP defines them; the mel points are FP64 values narrowed on the call.
struct MelScaleF32 {
    int formula;
    float hz_to_mel(float hz) const
    {
        if (formula == HTK) return 1127.0f * std::log(1.0f + (hz / 700.0f));
        const float fsp = 66.666667f, minLogHz = 1000.0f, minLogMel = (minLogHz - 0.0f) / fsp;
        const float stepLog = 0.068751777f;
        if (hz >= minLogHz) return minLogMel + std::log(hz * 0.001f) * (1.0f / stepLog);
        return (hz - 0.0f) * (1.0f / fsp);
    }
    float mel_to_hz(float mel) const
    {
        if (formula == HTK) return 700.0f * (std::exp(mel / 1127.0f) - 1.0f);
        const float fsp = 66.666667f, minLogHz = 1000.0f, minLogMel = (minLogHz - 0.0f) / fsp;
        const float stepLog = 0.068751777f;
        if (mel >= minLogMel) return minLogHz * std::exp(stepLog * (mel - minLogMel));
        return 0.0f + mel * fsp;
    }
};

// Filter corners of one bin count (cpu/kernel/mel_filter_bank.cpp:55-112): for each used bin,
// the filter whose rising slope it is on and its falling weight; each filter's area factor.
struct MelWeights {
    int binStart = 0, binEnd = 0;
    std::vector<float> down, norm;
    std::vector<int> up;
};

MelWeights mel_weights(int numBins, float sampleRate, float maxFreqVal, float minFreq, int formula, int M,
                       bool normalize)
{
    const MelScaleF32 sc{formula};
    const float maxFreq = maxFreqVal == 0 ? sampleRate / 2 : maxFreqVal;
    const double melLow = sc.hz_to_mel(minFreq), melHigh = sc.hz_to_mel(maxFreq);
    const double melStep = (melHigh - melLow) / (M + 1);
    const int nfft = (numBins - 1) * 2;
    const double hzStep = static_cast<double>(sampleRate) / nfft;
    MelWeights w;
    w.binStart = static_cast<int>(std::ceil(minFreq * (1.0 / hzStep)));
    w.binEnd = std::min(static_cast<int>(std::ceil(maxFreq * (1.0 / hzStep))), numBins);
    w.down.assign(numBins, 0.0f);
    w.up.assign(numBins, -1);
    w.norm.assign(M, 1.0f);
    int bin = w.binStart;
    double f = bin * hzStep;
    double m0 = melLow, m1 = melLow + melStep;
    for (int iv = 0; iv <= M; iv++, m0 = m1, m1 += melStep) {
        const double lo = sc.mel_to_hz(static_cast<float>(m0));
        const double hi = sc.mel_to_hz(static_cast<float>(iv == M ? melHigh : m1));
        const double inv = 1.0 / (hi - lo);
        if (normalize && iv < M) w.norm[iv] = static_cast<float>(2.0 / (sc.mel_to_hz(static_cast<float>(m1 + melStep)) - lo));
        for (; bin < w.binEnd && f < hi; bin++, f = bin * hzStep) {
            w.down[bin] = static_cast<float>((hi - f) * inv);
            w.up[bin] = iv;
        }
    }
    return w;
}

// RPP's reflect index over [0, L) (cpu/kernel/spectrogram.cpp:55-69), in closed form.
inline int reflect_index(int i, int L) { return idx_reflect(i, L); }

bool f32(const RpptDesc *d) { return d->dataType == F32; }

}  // namespace
}  // namespace host
}  // namespace rppa

using namespace rppa::host;

// cpu/kernel/pre_emphasis_filter.cpp:27-75: y[0] = x[0] - c*border (border x[0] CLAMP, x[1]
// REFLECT, 0 ZERO), y[t] = x[t] - c*x[t-1].  Clips of length 0 are left untouched (the reference
// reads x[0] and writes y[0] there).
RPPA_EXPORT RppStatus rppt_pre_emphasis_filter_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                                    RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor,
                                                    Rpp32f *coeffTensor, RpptAudioBorderType borderType,
                                                    rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (!f32(srcDescPtr) || !f32(dstDescPtr)) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcLengthTensor || !coeffTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    const float *src = static_cast<const float *>(srcPtr);
    float *dst = static_cast<float *>(dstPtr);
    const size_t sN = srcDescPtr->strides.nStride, dN = dstDescPtr->strides.nStride;
    for_each_clip(static_cast<int>(srcDescPtr->n), threads_of(rppHandle), [&](int b) {
        const float *x = src + b * sN;
        float *y = dst + b * dN;
        const int L = std::min<long long>(srcLengthTensor[b], static_cast<long long>(sN));
        if (L <= 0) return;
        const float c = coeffTensor[b];
        const float border = borderType == CLAMP ? x[0] : (borderType == REFLECT && L > 1 ? x[1] : 0.0f);
        y[0] = x[0] - c * border;
        for (int t = 1; t < L; t++) y[t] = x[t] - c * x[t - 1];
    });
    return RPP_SUCCESS;
}

// cpu/kernel/spectrogram.cpp:71-251.  Frames of W samples, hop H, centred (start f*H - W/2) or not,
// indices outside the clip reflected (without repeating the edge) or zero; each frame times the
// window (RPP's half-sample-shifted Hann unless windowFunction is given) sits at offset (N-W)/2 of
// an N-sample zero buffer; power == 1 gives |X|, any other value |X|^2 (the HOST rule,
// :217-242).  NFT output dst[k*hStride + f], NTF dst[f*hStride + k]; frames beyond the
// descriptor's frame capacity (w for NFT, h for NTF) are not written.  W > nfft is rejected (the
// reference writes past its FFT buffer there).
RPPA_EXPORT RppStatus rppt_spectrogram_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                            RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor, bool centerWindows,
                                            bool reflectPadding, Rpp32f *windowFunction, Rpp32s nfft, Rpp32s power,
                                            Rpp32s windowLength, Rpp32s windowStep, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (dstDescPtr->layout != NFT && dstDescPtr->layout != NTF) return RPP_ERROR_INVALID_DST_LAYOUT;
    if (!f32(srcDescPtr) || !f32(dstDescPtr)) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcLengthTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    if (nfft == 0) nfft = windowLength;
    if (windowLength <= 0 || windowStep <= 0 || nfft < 2 || windowLength > nfft) return RPP_ERROR_INVALID_ARGUMENTS;
    const bool nft = dstDescPtr->layout == NFT;
    const int K = nfft / 2 + 1, W = windowLength, H = windowStep, N = nfft;
    const int cap = static_cast<int>(nft ? dstDescPtr->w : dstDescPtr->h);
    const size_t hs = dstDescPtr->strides.hStride;
    if (nft ? hs < static_cast<size_t>(cap) : hs < static_cast<size_t>(K)) return RPP_ERROR_INVALID_DST_DIMS;
    std::vector<float> win(W);
    if (windowFunction) {
        std::memcpy(win.data(), windowFunction, sizeof(float) * W);
    } else {
        for (int t = 0; t < W; t++) win[t] = static_cast<float>(0.5 * (1.0 - std::cos(2.0 * M_PI * (t + 0.5) / W)));
    }
    const bool pow2 = N >= 4 && (N & (N - 1)) == 0;
    std::unique_ptr<RealFft> fft(pow2 ? new RealFft(N) : nullptr);
    std::unique_ptr<DirectDft> dft(pow2 ? nullptr : new DirectDft(N));
    const float *src = static_cast<const float *>(srcPtr);
    float *dst = static_cast<float *>(dstPtr);
    const size_t sN = srcDescPtr->strides.nStride, dN = dstDescPtr->strides.nStride;
    const int off = centerWindows ? W / 2 : 0;
    const int at = (N - W) / 2;
    for_each_clip(static_cast<int>(srcDescPtr->n), threads_of(rppHandle), [&](int b) {
        const float *x = src + b * sN;
        float *y = dst + b * dN;
        const int L = srcLengthTensor[b];
        const int F = std::min(frames_of(L, W, H, centerWindows), cap);
        std::vector<float> buf(N, 0.0f);
        std::vector<std::complex<float>> X(K), z(N / 2 + 1);
        for (int f = 0; f < F; f++) {
            const int s = f * H - off;
            float *fr = buf.data() + at;
            if (s >= 0 && s + W <= L) {
                for (int t = 0; t < W; t++) fr[t] = win[t] * x[s + t];
            } else if (reflectPadding) {
                for (int t = 0; t < W; t++) fr[t] = win[t] * x[reflect_index(s + t, L)];
            } else {
                for (int t = 0; t < W; t++) fr[t] = (s + t >= 0 && s + t < L) ? win[t] * x[s + t] : 0.0f;
            }
            if (fft) fft->run(buf.data(), X.data(), z.data());
            else dft->run(buf.data(), X.data());
            for (int k = 0; k < K; k++) {
                const float re = X[k].real(), im = X[k].imag();
                const float v = power == 1 ? std::abs(X[k]) : re * re + im * im;
                if (nft) y[k * hs + f] = v;
                else y[f * hs + k] = v;
            }
        }
    });
    return RPP_SUCCESS;
}

// cpu/kernel/mel_filter_bank.cpp:27-198.  srcDims[2b] = bins, srcDims[2b+1] = frames.  The first
// min(frames + 8, hStride) columns of every filter row are zeroed, then each bin in
// [ceil(fmin/hzStep), min(ceil(fmax/hzStep), bins)) adds (weight * power, rounded) to the filter
// below and the filter above it, bins ascending.
RPPA_EXPORT RppStatus rppt_mel_filter_bank_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                                RpptDescPtr dstDescPtr, Rpp32s *srcDims, Rpp32f maxFreq,
                                                Rpp32f minFreq, RpptMelScaleFormula melFormula, Rpp32s numFilter,
                                                Rpp32f sampleRate, bool normalize, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (srcDescPtr->layout != NFT) return RPP_ERROR_INVALID_SRC_LAYOUT;
    if (dstDescPtr->layout != NFT) return RPP_ERROR_INVALID_DST_LAYOUT;
    if (!f32(srcDescPtr) || !f32(dstDescPtr)) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcDims || numFilter <= 0) return RPP_ERROR_INVALID_ARGUMENTS;
    const float *src = static_cast<const float *>(srcPtr);
    float *dst = static_cast<float *>(dstPtr);
    const size_t sN = srcDescPtr->strides.nStride, sH = srcDescPtr->strides.hStride;
    const size_t dN = dstDescPtr->strides.nStride, dH = dstDescPtr->strides.hStride;
    const int M = numFilter;
    for_each_clip(static_cast<int>(srcDescPtr->n), threads_of(rppHandle), [&](int b) {
        const int bins = srcDims[2 * b], F = srcDims[2 * b + 1];
        float *y = dst + b * dN;
        const size_t zc = std::min<size_t>(static_cast<size_t>(std::max(F, 0)) + 8, dH);
        for (int m = 0; m < M; m++) std::fill(y + m * dH, y + m * dH + zc, 0.0f);
        if (bins < 2 || F <= 0) return;
        const MelWeights w = mel_weights(bins, sampleRate, maxFreq, minFreq, static_cast<int>(melFormula), M,
                                         normalize);
        const float *x = src + b * sN;
        for (int k = w.binStart; k < w.binEnd; k++) {
            const float *row = x + k * sH;
            const int up = w.up[k], down = up - 1;
            auto add = [&](int m, float wt) {
                float *o = y + m * dH;
                for (int f = 0; f < F; f++) {
                    const float p = row[f] * wt;
                    o[f] = o[f] + p;
                }
            };
            if (down >= 0) add(down, normalize ? w.down[k] * w.norm[down] : w.down[k]);
            if (up >= 0 && up < M) add(up, normalize ? (1.0f - w.down[k]) * w.norm[up] : 1.0f - w.down[k]);
        }
    });
    return RPP_SUCCESS;
}

// cpu/kernel/to_decibels.cpp:27-99: out = (multiplier * log10(2)) * log2(max(minRatio, x / ref)),
// minRatio = 10^(cutOffDB / multiplier) (the smallest positive float if that is 0), ref the
// given magnitude or, when it is 0, the clip's maximum (1 / ref read as 1 for a zero maximum).
// srcDims[b] = {width = columns, height = rows}; width 1 reads `height` contiguous values (1-D).
RPPA_EXPORT RppStatus rppt_to_decibels_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                            RpptDescPtr dstDescPtr, RpptImagePatchPtr srcDims, Rpp32f cutOffDB,
                                            Rpp32f multiplier, Rpp32f referenceMagnitude, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (!multiplier) return RPP_ERROR_ZERO_DIVISION;
    if (!f32(srcDescPtr) || !f32(dstDescPtr)) return RPP_ERROR_INVALID_SRC_OR_DST_DATATYPE;
    if (!srcPtr || !dstPtr || !srcDims) return RPP_ERROR_INVALID_ARGUMENTS;
    float minRatio = static_cast<float>(std::pow(10.0, static_cast<double>(cutOffDB / multiplier)));
    if (minRatio == 0.0f) minRatio = std::nextafter(0.0f, 1.0f);
    const float mult = multiplier * 0.3010299956639812f;
    const float *src = static_cast<const float *>(srcPtr);
    float *dst = static_cast<float *>(dstPtr);
    const size_t sN = srcDescPtr->strides.nStride, sH = srcDescPtr->strides.hStride;
    const size_t dN = dstDescPtr->strides.nStride, dH = dstDescPtr->strides.hStride;
    for_each_clip(static_cast<int>(srcDescPtr->n), threads_of(rppHandle), [&](int b) {
        const float *x = src + b * sN;
        float *y = dst + b * dN;
        const unsigned rows = srcDims[b].height, cols = srcDims[b].width;
        const bool oneD = cols == 1;
        float ref = referenceMagnitude;
        if (ref == 0.0f) {
            ref = -std::numeric_limits<float>::max();
            if (oneD) {
                for (unsigned i = 0; i < rows; i++) ref = std::max(ref, x[i]);
            } else {
                for (unsigned r = 0; r < rows; r++)
                    for (unsigned c = 0; c < cols; c++) ref = std::max(ref, x[r * sH + c]);
            }
        }
        const float inv = ref != 0.0f ? 1.f / ref : 1.0f;
        auto db = [&](float v) { return mult * std::log2(std::max(minRatio, v * inv)); };
        if (oneD) {
            for (unsigned i = 0; i < rows; i++) y[i] = db(x[i]);
        } else {
            for (unsigned r = 0; r < rows; r++)
                for (unsigned c = 0; c < cols; c++) y[r * dH + c] = db(x[r * sH + c]);
        }
    });
    return RPP_SUCCESS;
}

// cpu/kernel/down_mixing.cpp:28-122: mono[t] = sum over channels of x[t*C + c] * (1/C) (weights
// renormalised to sum 1 when asked), in the HOST kernel's grouping: channels in lanes of 4 (C < 8)
// or 8, each lane accumulating its channels in order, the lanes combined as the SSE/AVX horizontal
// adds combine them, the channels past the last full lane group added after.  srcDims[2b] =
// samples per channel, srcDims[2b+1] = channels; one channel is a copy.
RPPA_EXPORT RppStatus rppt_down_mixing_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                            RpptDescPtr dstDescPtr, Rpp32s *srcDimsTensor, bool normalizeWeights,
                                            rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (!f32(srcDescPtr) || !f32(dstDescPtr)) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcDimsTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    const float *src = static_cast<const float *>(srcPtr);
    float *dst = static_cast<float *>(dstPtr);
    const size_t sN = srcDescPtr->strides.nStride, dN = dstDescPtr->strides.nStride;
    for_each_clip(static_cast<int>(srcDescPtr->n), threads_of(rppHandle), [&](int b) {
        const float *x = src + b * sN;
        float *y = dst + b * dN;
        const long long S = srcDimsTensor[2 * b];
        const int C = srcDimsTensor[2 * b + 1];
        if (S <= 0 || C <= 0) return;
        if (C == 1) {
            std::memcpy(y, x, sizeof(float) * static_cast<size_t>(S));
            return;
        }
        std::vector<float> w(C, 1.f / C);
        if (normalizeWeights) {
            float sum = 0.0f;
            for (int c = 0; c < C; c++) sum += w[c];
            const float inv = static_cast<float>(1.0 / sum);
            for (int c = 0; c < C; c++) w[c] *= inv;
        }
        const int lanes = C >= 8 ? 8 : 4, full = C / lanes * lanes;
        for (long long t = 0; t < S; t++) {
            const float *v = x + t * C;
            float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int c0 = 0; c0 < full; c0 += lanes)
                for (int l = 0; l < lanes; l++) {
                    const float p = v[c0 + l] * w[c0 + l];
                    acc[l] = acc[l] + p;
                }
            float s;
            if (lanes == 4) {
                s = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            } else {
                const float a = acc[0] + acc[4], bb = acc[1] + acc[5], c = acc[2] + acc[6], d = acc[3] + acc[7];
                s = (a + c) + (bb + d);
            }
            for (int c = full; c < C; c++) {
                const float p = v[c] * w[c];
                s = s + p;
            }
            y[t] = s;
        }
    });
    return RPP_SUCCESS;
}

// cpu/kernel/resample.cpp:27-146 with the window of api/rppdefs.h:751-790: ceil(L * out / in)
// outputs in blocks of 256; per block an FP64 input position, then FP32 steps of in/out; taps
// [ceil(x) - lobes, ceil(x) + lobes) clipped to the clip; window value lookup[i] + frac *
// (lookup[i+1] - lookup[i]) at i = floor(u), u = tap offset * scale + center, clamped to the table
// (the 4-wide path of mono clips truncates and does not clamp).  Mono clips sum taps in four lanes
// (combined as (l0 + l2) + (l1 + l3)) and the remainder in order; several channels sum each channel
// tap by tap.  srcDims[2b] = samples per channel, srcDims[2b+1] = channels.  Equal rates copy.
RPPA_EXPORT RppStatus rppt_resample_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                         RpptDescPtr dstDescPtr, Rpp32f *inRateTensor, Rpp32f *outRateTensor,
                                         Rpp32s *srcDimsTensor, RpptResamplingWindow &window, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr || !dstDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (!f32(srcDescPtr) || !f32(dstDescPtr)) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !inRateTensor || !outRateTensor || !srcDimsTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    const float *lut = window.lookup.data();
    const int lutSize = window.lookupSize;
    if (!lut || lutSize < 2 || static_cast<size_t>(lutSize) > window.lookup.size()) return RPP_ERROR_INVALID_ARGUMENTS;
    const float wScale = window.scale, wCenter = window.center;
    const int lobes = window.lobes;
    auto weight = [&](float x) {
        const float u = x * wScale + wCenter;
        int i = static_cast<int>(std::floor(u));
        const float frac = u - i;
        i = std::max(std::min(i, lutSize - 2), 0);
        return lut[i] + frac * (lut[i + 1] - lut[i]);
    };
    auto weight_trunc = [&](float x) {
        const float u = x * wScale + wCenter;
        const int i = static_cast<int>(u);
        const float frac = u - static_cast<float>(i);
        return lut[i] + frac * (lut[i + 1] - lut[i]);
    };
    const float *src = static_cast<const float *>(srcPtr);
    float *dst = static_cast<float *>(dstPtr);
    const size_t sN = srcDescPtr->strides.nStride, dN = dstDescPtr->strides.nStride;
    for_each_clip(static_cast<int>(srcDescPtr->n), threads_of(rppHandle), [&](int b) {
        const float *x = src + b * sN;
        float *y = dst + b * dN;
        const float inRate = inRateTensor[b], outRate = outRateTensor[b];
        const int L = srcDimsTensor[2 * b], C = srcDimsTensor[2 * b + 1];
        if (L <= 0 || C <= 0) return;
        if (inRate == outRate) {
            std::memcpy(y, x, sizeof(float) * static_cast<size_t>(L) * C);
            return;
        }
        const int nOut = static_cast<int>(std::ceil(static_cast<float>(L) * outRate / inRate));
        const double step = static_cast<double>(inRate) / outRate;
        const float fstep = static_cast<float>(step);
        std::vector<float> acc(C);
        for (int o0 = 0; o0 < nOut; o0 += 256) {
            const int o1 = std::min(o0 + 256, nOut);
            const double base = o0 * step;
            const int baseIdx = static_cast<int>(base);
            float pos = static_cast<float>(base - baseIdx);
            const float *in = x + static_cast<size_t>(baseIdx) * C;
            for (int o = o0; o < o1; o++, pos += fstep) {
                const int xc = static_cast<int>(std::ceil(pos));
                const int t0 = std::max(xc - lobes, -baseIdx), t1 = std::min(xc + lobes, L - baseIdx);
                if (C == 1) {
                    float lane[4] = {0, 0, 0, 0};
                    const float first = static_cast<float>(t0) - pos;
                    float lp[4] = {first + 0.f, first + 1.f, first + 2.f, first + 3.f};
                    int t = t0;
                    for (; t + 3 < t1; t += 4)
                        for (int l = 0; l < 4; l++) {
                            const float p = in[t + l] * weight_trunc(lp[l]);
                            lane[l] = lane[l] + p;
                            lp[l] = lp[l] + 4.0f;
                        }
                    float s = (lane[0] + lane[2]) + (lane[1] + lane[3]);
                    float u = static_cast<float>(t) - pos;
                    for (; t < t1; t++, u += 1.0f) {
                        const float p = in[t] * weight(u);
                        s = s + p;
                    }
                    y[o] = s;
                } else {
                    std::fill(acc.begin(), acc.end(), 0.0f);
                    float u = static_cast<float>(t0) - pos;
                    for (int t = t0; t < t1; t++, u += 1.0f) {
                        const float wv = weight(u);
                        for (int c = 0; c < C; c++) {
                            const float p = in[static_cast<size_t>(t) * C + c] * wv;
                            acc[c] = acc[c] + p;
                        }
                    }
                    for (int c = 0; c < C; c++) y[static_cast<size_t>(o) * C + c] = acc[c];
                }
            }
        }
    });
    return RPP_SUCCESS;
}

// cpu/kernel/non_silent_region_detection.cpp:93-191: the moving mean square (running FP32 sum of
// squares over W samples, restarted from a fresh sum every resetInterval outputs, times 1/W),
// the first and last output at or above base * 10^(cutOffDB/10) (base = referencePower, or the
// largest mean square when that is 0), the begin moved W - 1 samples earlier (clamped at 0)
// unless it is 0.  A clip with no such output, or of length 0, gives (0, 0).
RPPA_EXPORT RppStatus rppt_non_silent_region_detection_host(RppPtr_t srcPtr, RpptDescPtr srcDescPtr,
                                                            Rpp32s *srcLengthTensor, Rpp32s *detectedIndexTensor,
                                                            Rpp32s *detectionLengthTensor, Rpp32f cutOffDB,
                                                            Rpp32s windowLength, Rpp32f referencePower,
                                                            Rpp32s resetInterval, rppHandle_t rppHandle)
{
    if (!rppHandle || !srcDescPtr) return RPP_ERROR_INVALID_ARGUMENTS;
    if (!f32(srcDescPtr)) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !srcLengthTensor || !detectedIndexTensor || !detectionLengthTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    if (windowLength < 1 || resetInterval == 0 || resetInterval < -1) return RPP_ERROR_INVALID_ARGUMENTS;
    const float cutOff = std::pow(10.0f, cutOffDB * 0.1f);
    const float *src = static_cast<const float *>(srcPtr);
    const size_t sN = srcDescPtr->strides.nStride;
    const int W = windowLength;
    for_each_clip(static_cast<int>(srcDescPtr->n), threads_of(rppHandle), [&](int b) {
        const float *x = src + b * sN;
        const int L = std::max(0, srcLengthTensor[b]);
        int begin = 0, length = 0;
        if (L > 0) {
            std::vector<float> mms(L);
            const int reset = resetInterval == -1 ? L : resetInterval;
            const float mf = 1.0f / W;
            int trail = 1 - W;   // the sample leaving the window after output p
            for (int p = 0; p < L;) {
                float sum = 0.0f;
                for (int i = std::max(trail, 0); i < p; i++) {
                    const float q = x[i] * x[i];
                    sum = sum + q;
                }
                const int end = std::min(L, p + reset);
                for (; p < end; p++, trail++) {
                    const float q = x[p] * x[p];
                    sum = sum + q;
                    mms[p] = sum * mf;
                    if (trail >= 0) {
                        const float r = x[trail] * x[trail];
                        sum = sum - r;
                    }
                }
            }
            const float base = referencePower == 0.0f ? *std::max_element(mms.begin(), mms.end()) : referencePower;
            const float cut = base * cutOff;
            int first = -1, last = -1;
            for (int i = 0; i < L; i++)
                if (mms[i] >= cut) { first = i; break; }
            if (first >= 0) {
                for (int i = L - 1; i >= first; i--)
                    if (mms[i] >= cut) { last = i; break; }
                begin = first;
                length = last - first + 1;
                if (begin != 0) {
                    const int nb = std::max(begin - (W - 1), 0);
                    length += begin - nb;
                    begin = nb;
                }
            }
        }
        detectedIndexTensor[b] = begin;
        detectionLengthTensor[b] = length;
    });
    return RPP_SUCCESS;
}
