// logmel_fft.hip -- the fused log-mel launch with the DFT factored into two MFMA stages.
//
// Same output contract as the folded dense kernel in logmel_fused.hip (pre-emphasis, framing with
// the reflect/zero rule, power spectrum, mel filter bank, dB over a ragged batch, one persistent
// launch), for nfft = 512 = N1 x N2 with N1 = 32 and N2 = 16, hops that are a multiple of 32 and
// windows up to nfft.  Buffer index n = 16 n1 + n2 of a frame (the window occupies n < W, the rest
// is zero; the origin only rotates the phase of each bin, which the power discards), bin
// k = k1 + 32 k2:
//
//   stage 1   Y[n2][k1] = sum_{n1 < 32} x[16 n1 + n2] w[16 n1 + n2] W32^(n1 k1)      (real input)
//   twiddle   T[n2][k1] = W512^(n2 k1) Y[n2][k1]                                   (k1 = 1 .. 15)
//   stage 2   X[k1 + 32 k2] = sum_{n2 < 16} W16^(n2 k2) T[n2][k1]
//
// Real input makes Y[n2][32 - k1] = conj Y[n2][k1], so k1 = 0 .. 16 suffice: for k1 = 1 .. 15 the
// sixteen k2 outputs are eight bins k1 + 32 k2 (k2 < 8) and, conjugated, eight bins
// (32 - k1) + 32 (15 - k2) (k2 >= 8).  The bins k = 16 q (k1 = 0 and 16) form one more product: with
// Y[n2][0] = sum x and Y[n2][16] = sum (-1)^n1 x, X[16 q] = sum_n2 W32^(q n2) Y[n2][q odd ? 16 : 0].
// Bin N/2 is never read: the mel filters stop below it (maxFreq <= sampleRate / 2).
//
// Per dense-equivalent frame this is 2 x (16 x 32 x 32) + 2 x (16 x 32 x 32) = 65,536 MFMA FLOP per
// product pass instead of 2 x 400 x 514 = 411,200 for the dense DFT.  Both stages run on
// v_mfma_f32_16x16x32_f16 with split hi+lo operands (hi*hi + hi*lo + lo*hi, FP32 accumulate):
// the samples are scaled per work item by a power of two so |x w| < 2^10, the DFT matrices by 2^12,
// the stage-1 outputs by 2^-12 (folded into the twiddles) before their split; every scaling is
// exact and undone once per mel output.
//
// Work decomposition: one wave owns one work item (16 consecutive frames of one clip) at a time,
// with its own LDS segment, so the item loop has no workgroup barrier.  Lane l works on frame
// l % 16 and lane group g = l / 16.
//   staging   the item's pre-emphasised samples, scaled, in a polyphase layout seg[n2][pos] (sample
//             16 pos + n2 of the segment): frame f's inputs of phase n2 are the consecutive positions
//             f H/16 + n1, read as 8-byte pairs (n1 = 8 g .. 8 g + 7 per lane);
//   stage 1   per n2: the lane forms x w for its eight n1 and splits them (one v_fma_mix per half,
//             each rounding the exact product once), then 2 row tiles x 3 products against the
//             32-point real DFT matrix (rows: re/im of k1, with k1 = 0 carrying Y[.][0] and Y[.][16]);
//             lane group g ends up holding rows 4g .. 4g+3 of both row tiles, i.e. k1 = 2g, 2g+1,
//             8+2g, 9+2g, for all sixteen n2;
//   twiddle   in FP32, then split into packed fp16 pairs of consecutive n2;
//   transpose the stage-2 operand of k1 needs (re, im) x n2 spread over the four lane groups (group
//             g' holds contraction elements 8g' .. 8g'+7): a 4 x 4 transpose across lane groups per
//             register, two rounds of v_permlane32_swap / v_permlane16_swap (one instruction per
//             register) -- no LDS round trip;
//   stage 2   per k1: 2 row tiles x 3 products; the rows are ordered so lane group g holds bins
//             [64g, 64g + 64) of its frame; power in lane;
//   mel       the power values split into bf16 hi/lo pairs and transposed the same way, so a lane
//             group holds eight consecutive bins of every 32-bin chunk; banded mel^T = W_mel . P on
//             v_mfma_f32_16x16x32_bf16 (hi*hi + hi*lo + lo*hi) over the (16-filter tile, chunk) pairs
//             the filters touch, the weight fragments in LDS;
//   dB        in registers, stored as 64-byte runs of a mel row (NFT) or four filters (NTF).
#include "common.hpp"
#include "down_mix_math.hpp"
#include "resample_math.hpp"

#include <cmath>

namespace rppa {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr double kPi = 3.14159265358979323846;
constexpr int kN1 = 32, kN2 = 16;     // nfft 512 = 32 x 16
constexpr int kN = kN1 * kN2;
constexpr int kFI = 16;               // frames per work item
constexpr int kWaves = 8;             // waves per workgroup (two per SIMD)
constexpr int kThreads = 64 * kWaves;
constexpr int kWavesR2 = 8;           // nfft 1024: waves per workgroup where the LDS holds eight segments,
constexpr int kWavesR2Min = 6;        // else six
constexpr int kPitchR2 = 88;          // nfft 1024: floats per phase row (odd rows skewed by 4 more)
constexpr int kMaxMT = 8;             // filter tiles of 16 (numFilter <= 128)
constexpr int kChunks = kN / 2 / 32;  // 32-bin chunks of bins 0 .. N/2 - 1
// the mel window of a launch with NMT filter tiles: the filter tiles one 32-bin chunk may touch (all
// of them when NMT <= 5, else at most six)
constexpr int mel_window(int nmt) { return nmt < 6 ? nmt : 6; }
// fused front end: 2 x 16 FIR taps (16 zero crossings), resampled outputs per chunk (two per lane)
// and the down-mixed inputs a chunk reads at up to 3 inputs per output
constexpr int kFeTaps = 32;
constexpr int kFeOut = 128;
constexpr int kFeMaxK = 3;
constexpr int kFeSpan = 448;
static_assert(kFeSpan >= kFeMaxK * (kFeOut - 1) + kFeTaps && kFeSpan % 64 == 0, "front-end span");

// ------------------------------------------------------------------------- tables (FP64)
__device__ inline uint16_t f16_bits(double v)
{
    _Float16 h = static_cast<_Float16>(static_cast<float>(v));
    uint16_t r;
    __builtin_memcpy(&r, &h, 2);
    return r;
}
__device__ inline double f16_value(uint16_t u)
{
    _Float16 h;
    __builtin_memcpy(&h, &u, 2);
    return static_cast<double>(static_cast<float>(h));
}
// v scaled by 2^12, split into fp16 hi + lo (each rounded to nearest even)
__device__ inline void split_f16_scaled(double v, uint16_t &hi, uint16_t &lo)
{
    const double s = v * 4096.0;
    hi = f16_bits(s);
    lo = f16_bits(s - f16_value(hi));
}

// One thread per table element.  Fragment tables are in MFMA operand lane order: [..][hi/lo][lane 64][8]
// with lane l holding row l % 16 and contraction elements 8 (l / 16) .. + 7.
//   a1  [rt 2][hl][64][8]  stage-1 matrix: row r = 16 rt + 4 gg + 2 pair + part is (re, im) of
//                          k1 = 8 rt + 2 gg + pair (k1 = 0: Y[.][0] and Y[.][16]); column n1
//   a2  [which 2][ort 2][hl][64][8]  stage 2, contraction c = 16 part + n2:
//                          which 0: the 16-point DFT; row 16 ort + 2 m + po is part po of
//                          k2 = m (ort 0) or 15 - m (ort 1)
//                          which 1: bins 16 q from (Y[.][0], Y[.][16]); row 16 ort + 4 gg + 2 qq + po is
//                          part po of q = 4 gg + 2 ort + qq
//   winP [n2][n1]          w[16 n1 + n2] (zero at and past W); nfft 1024: [p 2][n2][n1] w[32 n1 + 2 n2 + p]
//   tw   [k1 16][n2 16][2] 2^-12 (cos, sin)(2 pi n2 k1 / 512); k1 = 0: (2^-12, 0)
//   tw2  nfft 1024 only, [g 4][k1 16][j 4][2]: for value j of the stage-2 slot k1 in lane group g (bin
//        k = 64 g + (k1 ? {k1, k1 + 32, 32 - k1, 64 - k1}[j] : 16 j)) the parity-0 factor
//        conj(W1024^k) = (cos, sin)(2 pi k / 1024), conjugated for the values the slot holds
//        conjugated (k1 > 0, j >= 2)
__global__ void fft_tables_kernel(int W, int nfft, const float *__restrict__ win, uint16_t *__restrict__ a1,
                                  uint16_t *__restrict__ a2, float *__restrict__ winP, float *__restrict__ tw,
                                  float *__restrict__ tw2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // a1: 2 rt x 512 elements per plane
    if (t < 2 * 512) {
        const int rt = t >> 9, lane = (t >> 3) & 63, j = t & 7;
        const int rr = lane & 15, n1 = 8 * (lane >> 4) + j;
        const int gg = rr >> 2, pair = (rr >> 1) & 1, part = rr & 1;
        const int k1 = 8 * rt + 2 * gg + pair;
        double v;
        if (k1 == 0) {
            v = part == 0 ? 1.0 : ((n1 & 1) ? -1.0 : 1.0);
        } else {
            const double th = 2.0 * kPi * static_cast<double>((n1 * k1) % kN1) / kN1;
            v = part == 0 ? cos(th) : -sin(th);
        }
        uint16_t hi, lo;
        split_f16_scaled(v, hi, lo);
        a1[((rt * 2 + 0) * 64 + lane) * 8 + j] = hi;
        a1[((rt * 2 + 1) * 64 + lane) * 8 + j] = lo;
    }
    // a2: 2 which x 2 ort x 512 elements per plane
    if (t < 4 * 512) {
        const int which = t >> 10, ort = (t >> 9) & 1, lane = (t >> 3) & 63, j = t & 7;
        const int rr = lane & 15, c = 8 * (lane >> 4) + j;
        const int partIn = c >> 4, n2 = c & 15;
        double v = 0.0;
        if (which == 0) {
            const int m = rr >> 1, po = rr & 1;
            const int k2 = ort == 0 ? m : 15 - m;
            const double th = 2.0 * kPi * static_cast<double>((n2 * k2) % kN2) / kN2;
            if (po == 0) v = partIn == 0 ? cos(th) : sin(th);
            else v = partIn == 0 ? -sin(th) : cos(th);
        } else {
            const int gg = rr >> 2, qq = (rr >> 1) & 1, po = rr & 1;
            const int q = 4 * gg + 2 * ort + qq;
            if ((q & 1) == partIn) {
                const double ph = 2.0 * kPi * static_cast<double>((q * n2) % 32) / 32.0;
                v = po == 0 ? cos(ph) : -sin(ph);
            }
        }
        uint16_t hi, lo;
        split_f16_scaled(v, hi, lo);
        a2[(((which * 2 + ort) * 2 + 0) * 64 + lane) * 8 + j] = hi;
        a2[(((which * 2 + ort) * 2 + 1) * 64 + lane) * 8 + j] = lo;
    }
    if (nfft == 2 * kN) {
        if (t < 2 * kN2 * kN1) {
            const int p = t / (kN2 * kN1), n2 = (t / kN1) % kN2, n1 = t % kN1, n = 32 * n1 + 2 * n2 + p;
            winP[t] = n < W ? win[n] : 0.0f;
        }
        if (t < 4 * 16 * 4) {
            const int g = t >> 6, k1 = (t >> 2) & 15, j = t & 3;
            const int idx = k1 == 0 ? 16 * j : (j == 0 ? k1 : (j == 1 ? k1 + 32 : (j == 2 ? 32 - k1 : 64 - k1)));
            const double th = 2.0 * kPi * static_cast<double>(64 * g + idx) / (2 * kN);
            const double sgn = (k1 > 0 && j >= 2) ? -1.0 : 1.0;
            tw2[2 * t] = static_cast<float>(cos(th));
            tw2[2 * t + 1] = static_cast<float>(sgn * sin(th));
        }
    } else if (t < kN2 * kN1) {
        const int n2 = t / kN1, n1 = t % kN1, n = 16 * n1 + n2;
        winP[t] = n < W ? win[n] : 0.0f;
    }
    if (t < 16 * 16) {
        const int k1 = t >> 4, n2 = t & 15;
        double c = 1.0, s = 0.0;
        if (k1 > 0) {
            const double th = 2.0 * kPi * static_cast<double>((n2 * k1) % kN) / kN;
            c = cos(th);
            s = sin(th);
        }
        tw[2 * t] = static_cast<float>(c / 4096.0);
        tw[2 * t + 1] = static_cast<float>(s / 4096.0);
    }
}

// RPP's Hann (cpu/kernel/spectrogram.cpp:36-44): FP64, rounded to FP32.
__global__ void fft_hann_kernel(int W, float *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < W) out[t] = static_cast<float>(0.5 * (1.0 - cos((2.0 * kPi / W) * (t + 0.5))));
}

__device__ inline uint16_t bf16_bits(float v)
{
    const uint32_t u = __float_as_uint(v);
    return static_cast<uint16_t>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Mel weight fragments of the (16-filter tile, 32-bin chunk) pairs the filters touch, bf16 hi/lo in
// A-operand lane order ([pair][hl][64][8]: lane l holds filter 16 mt + l % 16, positions
// 32 c + 8 (l / 16) .. + 7 of the chunk).  Pairs are numbered chunk by chunk, tiles ascending within
// a chunk: the tiles whose band [c0, c1) contains chunk c are a contiguous run [t0, t1) (the filters
// ascend in frequency), so pairOff[c] = t0 | t1 << 8 | p0 << 16 with p0 the run's first pair (t0 = t1:
// no pair), pairOff[256] = pair count, pairOff[255] = 1 if some chunk's tiles are not contiguous, a
// run touches more tiles than the mel window holds or the runs' first tiles descend (the factored
// kernel then cannot use the table).  One workgroup.
// r2 = 0: position u is bin u (bins 0 .. N/2 - 1).  r2 = 1 (nfft 1024, radix-2 split): each chunk has
// two runs, first the lower half's (position u is bin u < 256, word pairOff[c]) and then the upper
// half's (position u is bin 512 - u, bins 257 .. 512, word pairOff[16 + c]), numbered in that order;
// bin 256's weights go to w256[M].
__global__ void __launch_bounds__(256) mel_frag_kernel(const float *__restrict__ dense, const int *__restrict__ filterBins,
                                                       int M, int numBins, int r2, uint16_t *__restrict__ frag,
                                                       int *__restrict__ pairOff, float *__restrict__ w256)
{
    __shared__ int band[2][16][2];
    __shared__ int pairOf[2][16][16];   // [half][chunk][tile]
    __shared__ int word[32];
    __shared__ int total, broken;
    const int nMT = (M + 15) / 16;
    const int nCh = r2 ? kChunks : min((numBins - 1) / 32, kChunks);   // bin N/2 carries no weight
    const int halves = r2 ? 2 : 1;
    const int tid = threadIdx.x;
    if (tid < 32) {
        const int half = tid >> 4, mt = tid & 15;
        int lo = 1 << 30, hi = 0;
        for (int i = 0; i < 16 && mt < nMT; i++) {
            const int m = mt * 16 + i;
            if (m < M && filterBins[2 * m] < filterBins[2 * m + 1]) {
                lo = min(lo, filterBins[2 * m]);
                hi = max(hi, filterBins[2 * m + 1]);
            }
        }
        if (r2 && half == 0) hi = min(hi, 256);
        if (r2 && half == 1) {
            // bins [max(lo, 257), min(hi, 513)) are positions u = 512 - bin in [513 - hi', 513 - lo')
            const int bl = max(lo, 257), bh = min(hi, 513);
            lo = bl < bh ? 513 - bh : 0;
            hi = bl < bh ? 513 - bl : 0;
        }
        band[half][mt][0] = lo < hi ? lo / 32 : 0;
        band[half][mt][1] = lo < hi ? min((hi + 31) / 32, nCh) : 0;
    }
    if (r2)
        for (int m = tid; m < M; m += blockDim.x) w256[m] = dense[static_cast<size_t>(m) * numBins + 256];
    __syncthreads();
    if (tid == 0) {
        int n = 0, bad = 0;
        for (int c = 0; c < 16; c++)
            for (int half = 0; half < 2; half++) {
                int t0 = -1, t1 = -1;
                for (int mt = 0; mt < 16; mt++) {
                    const bool in = half < halves && mt < nMT && c >= band[half][mt][0] && c < band[half][mt][1];
                    pairOf[half][c][mt] = -1;
                    if (!in) continue;
                    if (t0 < 0) t0 = mt;
                    else if (t1 != mt) bad = 1;   // a gap in the run
                    t1 = mt + 1;
                    pairOf[half][c][mt] = n + (mt - t0);
                }
                if (t0 < 0) t0 = t1 = 0;
                word[16 * half + c] = t0 | (t1 << 8) | (n << 16);
                n += t1 - t0;
            }
        // the kernel streams the runs in ascending frequency (the lower half's by ascending chunk,
        // then the upper half's by descending chunk) with the mel window's tiles in flight (the
        // launch's tile count: 5 up to 80 filters, else 8): each run's first tile must not decrease,
        // and its tiles must fit the window
        const int win = mel_window(nMT <= 5 ? 5 : 8);
        for (int r = 0, prev = 0; r < halves * kChunks; r++) {
            const int w = word[r < kChunks ? r : 16 + (2 * kChunks - 1 - r)];
            const int t0 = w & 255, t1 = (w >> 8) & 255;
            if (t0 >= t1) continue;
            if (t0 < prev || t1 - t0 > win) bad = 1;
            prev = t0;
        }
        total = n;
        broken = bad;
    }
    __syncthreads();
    for (int i = tid; i < 257; i += blockDim.x)
        pairOff[i] = i < 16 * halves ? word[i] : (i == 256 ? total : (i == 255 ? broken : 0));
    for (int hcm = 0; hcm < 512; hcm++) {
        const int half = hcm >> 8, c = (hcm >> 4) & 15, mt = hcm & 15, p = pairOf[half][c][mt];
        if (p < 0) continue;
        for (int e = tid; e < 512; e += blockDim.x) {
            const int lane = e >> 3, j = e & 7;
            const int m = mt * 16 + (lane & 15), u = 32 * c + 8 * (lane >> 4) + j;
            const int k = half ? 512 - u : u;
            const float v = (m < M && k < numBins) ? dense[static_cast<size_t>(m) * numBins + k] : 0.0f;
            const uint16_t h = bf16_bits(v);
            const uint16_t l = bf16_bits(v - __uint_as_float(static_cast<uint32_t>(h) << 16));
            frag[((static_cast<size_t>(p) * 2 + 0) * 64 + lane) * 8 + j] = h;
            frag[((static_cast<size_t>(p) * 2 + 1) * 64 + lane) * 8 + j] = l;
        }
    }
}

// ------------------------------------------------------------------------- kernel
struct FftArgs {
    const float *src;
    uint32_t srcStride;
    const int *lengths;
    const int *begin;
    int B;
    const long long *srcOff, *dstOff;
    float *dst;
    uint32_t dstNStride, hStride;
    int ntf;
    uint32_t frameCap;
    int *clipMax;
    float coeff;
    int border;
    int W, H, center, reflect, power, M;
    float minRatio, mult, invRef;
    const uint16_t *a1, *a2;
    const float *winP, *tw;
    const uint16_t *melFrag;
    const int *pairOff;
    int nPairs;
    // nfft 1024 (radix-2 launch): the radix-2 twiddles and the mel weights of bin 256
    const float *tw2, *w256;
    int pitch;       // floats per phase row of a wave's segment
    // fused front end (FRONT kernels): interleaved multi-channel clips fsrc[b * fStride + t * C + c],
    // fDims[2 b] samples per channel and fDims[2 b + 1] channels, down-mixed and resampled by the
    // integer ratio fK = inRate / outRate with the window table fLut; the resampled clip is what the
    // rest of the kernel sees
    const float *fsrc;
    uint32_t fStride;
    const int *fDims;
    int fK, fNormalize;
    float fIr, fOr;
    const float *fLut;
    float fWScale, fWCenter;
    int fLookup;
};

struct Clip {
    const float *x;
    float *out;
    uint32_t hs;
    int L, F, b;
};

// Clip b from the per-clip metadata staged in LDS (length, region start or packed offsets); every
// value is wave-uniform.
__device__ __forceinline__ Clip clip_of(const FftArgs &a, int b, const int *lenL, const int *startL,
                                        const long long *offL)
{
    Clip v;
    v.L = __builtin_amdgcn_readfirstlane(lenL[b]);
    long long xo, oo;
    if (a.srcOff) {
        auto uniform64 = [](long long u) {
            const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(u));
            const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(u >> 32));
            return (static_cast<long long>(hi) << 32) | static_cast<uint32_t>(lo);
        };
        xo = uniform64(offL[2 * b]);
        oo = uniform64(offL[2 * b + 1]);
    } else {
        xo = static_cast<long long>(b) * a.srcStride + (a.begin ? __builtin_amdgcn_readfirstlane(startL[b]) : 0);
        oo = static_cast<long long>(b) * a.dstNStride;
    }
    v.x = a.src + xo;
    int F = max(num_windows(v.L, a.W, a.H, a.center != 0), 0);
    v.out = a.dst + oo;
    if (a.dstOff) {
        v.hs = static_cast<uint32_t>(F);
    } else {
        v.hs = a.hStride;
        F = min(F, static_cast<int>(a.frameCap));
    }
    v.F = F;
    v.b = b;
    return v;
}

__device__ __forceinline__ int items_of(const FftArgs &a, int L, int FI)
{
    int F = max(num_windows(L, a.W, a.H, a.center != 0), 0);
    if (!a.dstOff) F = min(F, static_cast<int>(a.frameCap));
    return (F + FI - 1) / FI;
}

// The value of v in lane l ^ 8 (the other half of the lane's 16-lane row): DPP row rotate by 8.
__device__ __forceinline__ uint32_t swap8u(uint32_t v)
{
    return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x128, 0xf, 0xf, false));
}
__device__ __forceinline__ float swap8(float v) { return __uint_as_float(swap8u(__float_as_uint(v))); }

// Pre-emphasised sample j of a clip of length L with the spectrogram's padding rule (the chain pads
// the pre-emphasised signal): the reflect map for indices outside the clip, zero without reflect;
// the border rule at index 0.  The product is rounded before the subtraction, as RPP HOST rounds it.
__device__ __forceinline__ float pe_at(const float *x, int L, int j, float c, int border, int reflect)
{
    if (L <= 0) return 0.0f;
    if (j < 0 || j >= L) {
        if (!reflect) return 0.0f;
        j = idx_reflect(j, L);
    }
    float prev;
    if (j == 0) prev = (border == CLAMP) ? x[0] : (border == REFLECT && L > 1 ? x[1] : 0.0f);
    else prev = x[j - 1];
    const float p = c * prev;
    return x[j] - p;
}

// RPP's reflect index (idx_reflect) with the common single bounce done without a division.
__device__ __forceinline__ int reflect_fast(int j, int L)
{
    if (L < 2) return L - 1;
    int r = j < 0 ? -j : 2 * (L - 1) - j;
    if (r < 0 || r >= L) r = idx_reflect(j, L);
    return r;
}

// The value of v, hidden from the optimiser: lane-index arithmetic derived from it inside the item
// loop stays in its phase instead of being hoisted into registers that live across the whole loop.
__device__ __forceinline__ int opaque(int v)
{
    __asm__ volatile("" : "+v"(v));
    return v;
}

// One work item: its clip, first frame, segment start and the samples that hold frame data.
struct Item {
    Clip v;
    int f0, s0, valid;
    bool interior;   // every staged sample and its predecessor inside the clip, valid a multiple of 4
};

// fp16 hi/lo of the exact products d0*w0 and d1*w1 as two packed pairs: each half is one
// v_fma_mix, hi = f16(d w), lo = f16(d w - hi), each rounded once.
__device__ __forceinline__ void split_mul(float d0, float w0, float d1, float w1, uint32_t &hi, uint32_t &lo)
{
    uint32_t h, l;
    __asm__("v_fma_mixlo_f16 %0, %1, %2, 0" : "=v"(h) : "v"(d0), "v"(w0));
    __asm__("v_fma_mixhi_f16 %0, %1, %2, 0" : "+v"(h) : "v"(d1), "v"(w1));
    __asm__("v_fma_mixlo_f16 %0, %1, %2, -%3 op_sel_hi:[0,0,1]" : "=v"(l) : "v"(d0), "v"(w0), "v"(h));
    __asm__("v_fma_mixhi_f16 %0, %1, %2, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(l) : "v"(d1), "v"(w1), "v"(h));
    hi = h;
    lo = l;
}

// fp16 hi/lo of two FP32 values: hi = f16(v) (one packed conversion, round to nearest even),
// lo = f16(v - hi) (one v_fma_mix per half: the difference is formed exactly, rounded once).
__device__ __forceinline__ void split_pair(float v0, float v1, uint32_t &hi, uint32_t &lo)
{
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
    const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{v0, v1}, f16x2));
    uint32_t l;
    __asm__("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(l) : "v"(v0), "v"(h));
    __asm__("v_fma_mixhi_f16 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(l) : "v"(v1), "v"(h));
    hi = h;
    lo = l;
}

__device__ __forceinline__ f32x4 mfma3_f16(const u32x4 ah, const u32x4 al, const u32x4 bh, const u32x4 bl, f32x4 acc)
{
    const f16x8 a0 = __builtin_bit_cast(f16x8, ah), a1 = __builtin_bit_cast(f16x8, al);
    const f16x8 b0 = __builtin_bit_cast(f16x8, bh), b1 = __builtin_bit_cast(f16x8, bl);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b1, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, b0, acc, 0, 0, 0);
    return acc;
}

__device__ __forceinline__ f32x4 mfma3_bf16(const u32x4 ah, const u32x4 al, const u32x4 bh, const u32x4 bl, f32x4 acc)
{
    const bf16x8 a0 = __builtin_bit_cast(bf16x8, ah), a1 = __builtin_bit_cast(bf16x8, al);
    const bf16x8 b0 = __builtin_bit_cast(bf16x8, bh), b1 = __builtin_bit_cast(bf16x8, bl);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc, 0, 0, 0);
    return acc;
}

// 4 x 4 transpose across the lane groups (rows of 16 lanes): on entry r[v] holds piece (v, g) in
// lane group g, on exit piece (g, v).  permlane32_swap exchanges rows 2-3 of its first operand with
// rows 0-1 of its second; permlane16_swap exchanges the odd rows of its first operand with the even
// rows of its second.
__device__ __forceinline__ void transpose4(uint32_t &r0, uint32_t &r1, uint32_t &r2, uint32_t &r3)
{
    auto s02 = __builtin_amdgcn_permlane32_swap(r0, r2, false, false);
    auto s13 = __builtin_amdgcn_permlane32_swap(r1, r3, false, false);
    auto t01 = __builtin_amdgcn_permlane16_swap(s02[0], s13[0], false, false);
    auto t23 = __builtin_amdgcn_permlane16_swap(s02[1], s13[1], false, false);
    r0 = t01[0];
    r1 = t01[1];
    r2 = t23[0];
    r3 = t23[1];
}

// m = max(m, |a|, |b|) in one instruction (the library max would first canonicalise each loaded
// operand; the magnitudes here only set a power-of-two scale)
__device__ __forceinline__ void max3_abs(float &m, float a, float b)
{
    __asm__("v_max3_f32 %0, |%1|, |%2|, %0" : "+v"(m) : "v"(a), "v"(b));
}

// The wave's maximum of non-negative values, uniform: DPP steps within each 16-lane row (lane swaps
// in quads, half-row and row mirrors), then row broadcasts carry the rows' maxima up to lane 63.  Lanes
// a step has no source for take 0, which a non-negative maximum ignores.  Vector ALU only: no LDS
// permute, whose latency each step would wait out.
__device__ __forceinline__ float wave_max_nonneg(float v)
{
#define RPPA_DPP_MAX(CTRL, ROWS) \
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false)))
    RPPA_DPP_MAX(0xb1, 0xf);    // quad_perm [1, 0, 3, 2]
    RPPA_DPP_MAX(0x4e, 0xf);    // quad_perm [2, 3, 0, 1]
    RPPA_DPP_MAX(0x141, 0xf);   // row_half_mirror
    RPPA_DPP_MAX(0x140, 0xf);   // row_mirror: every lane holds its row's maximum
    RPPA_DPP_MAX(0x142, 0xa);   // row_bcast:15 into rows 1 and 3
    RPPA_DPP_MAX(0x143, 0xc);   // row_bcast:31 into rows 2 and 3
#undef RPPA_DPP_MAX
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Two FP32 values -> packed bf16 hi pair and packed bf16 lo pair (lo = bf16(v - hi)).
__device__ __forceinline__ void split_bf16_pair(float p0, float p1, uint32_t &hi, uint32_t &lo)
{
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const uint32_t hp = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{p0, p1}, bf16x2));
    const float h0 = __uint_as_float(hp << 16), h1 = __uint_as_float(hp & 0xffff0000u);
    hi = hp;
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{p0 - h0, p1 - h1}, bf16x2));
}

// Phase-row pitch of a wave's segment for `positions` used positions: a multiple of 4 (16-byte row
// starts) with pitch % 32 == 28, so the four phase rows one staging store touches fall on two bank
// sets.
constexpr int pitch_for(int positions)
{
    int p = (positions + 3) / 4 * 4;
    while (p % 32 != 28) p += 4;
    return p;
}

// GENERAL = false: the common launch (power 2, NFT output, a fixed dB reference, a cut-off whose
// ratio is a normal float); true: any power, NTF, the clip-maximum reference, subnormal ratios.
// HOP: the hop in samples (a multiple of 32), which fixes the segment geometry.
// FRONT: 0, or the clips arrive interleaved at FRONT (2 or 3) times the rate and each item computes
// the down-mixed, resampled samples it reads (bit for bit what the down-mix + resample launch
// writes) instead of loading them.
//
// NFFT 1024 (R2): radix-2 over the same 512-point machinery.  A 1024-point frame's even and odd
// samples are two columns of the 512-point DFT (column l % 16 = frame l % 8, parity (l / 8) % 2;
// 8 frames per item, the segment in 32 polyphase rows so column (f, p) reads rows 2 n2 + p); after
// stage 2 the lanes of a frame's two parities swap their values (lanes l and l ^ 8) and each forms
// X[k] = E[k] + W1024^k O[k] (parity 0, bin k) or X[512 - k] = conj(E[k] - W1024^k O[k]) (parity 1,
// bin 512 - k) for the k < 256 it holds; bin 256 is E[256]^2 + O[256]^2 from the stage-1 sums.  The
// mel runs over both halves with their own fragment tables (the lower half's columns valid in the
// parity-0 accumulators, the upper half's in the parity-1 ones), which are then added across the
// lane pair together with bin 256's weights.
template <int NMT, int HOP, bool GENERAL, int FRONT, int NFFT = 512, int WAVES = kWaves>
__global__ void __launch_bounds__(64 * WAVES, 1) logmel_fft_kernel(FftArgs a)
{
    constexpr bool R2 = NFFT == 1024;
    constexpr int FI = R2 ? 8 : kFI;                 // frames per item
    constexpr int NW = WAVES;                        // waves per workgroup
    constexpr int NT = 64 * NW;
    constexpr int NPH = R2 ? 32 : 16;                // polyphase rows of a segment
    constexpr int HQ = HOP / NPH;                    // positions per hop
    constexpr int SEGN = (FI - 1) * HOP + NFFT;      // segment samples (zero past the frames' data)
    // phase row r of a segment starts at row_at(r): nfft 512 a uniform pitch; nfft 1024 a pitch of
    // 88 with every odd row 4 floats further on, so that a frame's two parities (rows 2 n2 and
    // 2 n2 + 1) fall on different LDS banks
    constexpr int PITCH = R2 ? kPitchR2 : pitch_for(SEGN / NPH);
    constexpr int SEGF = R2 ? NPH * PITCH + 4 * (NPH / 2) : NPH * PITCH;   // floats per segment
    auto row_at = [](int r) { return R2 ? r * PITCH + 4 * ((r + 1) >> 1) : r * PITCH; };
    static_assert(!R2 || (FRONT == 0 && !GENERAL), "the radix-2 launch is the common one");
    static_assert(!R2 || PITCH >= SEGN / NPH, "nfft 1024 segment rows");
    constexpr int ROUNDS = (SEGN + 1023) / 1024;
    // staging rounds requested one item ahead (nfft 1024: the first only -- its stage-2 combine leaves
    // no registers to hold the others across the mel phase; they are requested at staging)
    constexpr int PFR = R2 ? 1 : ROUNDS;
    static_assert(HOP % 32 == 0 && ROUNDS <= 3, "segment geometry");
    static_assert(!FRONT || PITCH - SEGN / NPH >= 2, "two spare columns for the front end's predecessors");
    extern __shared__ __attribute__((aligned(16))) float lds[];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    // LDS: stage fragments a1 (4 KB) and a2 (8 KB), window [16][32], twiddles [16][16] float2, mel
    // fragments, tile plan, the waves' segments
    u32x4 *A1 = reinterpret_cast<u32x4 *>(lds);
    u32x4 *A2 = A1 + 256;
    float *winP = reinterpret_cast<float *>(A2 + 512);
    float *tw = winP + (R2 ? 2 : 1) * kN2 * kN1;
    const float4 *tw2 = reinterpret_cast<const float4 *>(tw + 2 * 256);   // R2: radix-2 twiddles [g][k1][4]
    float *w256 = tw + 2 * 256 + (R2 ? 2 * 256 : 0);                      // R2: bin 256's mel weights
    u32x4 *melF = reinterpret_cast<u32x4 *>(w256 + (R2 ? 16 * NMT : 0));
    // per clip (B rounded up to 4): the plan ts[0 .. B], the region length, the region start (with a
    // begin array) or the packed offsets; then the mel chunk words, then the waves' segments
    const int B4 = (a.B + 1 + 3) / 4 * 4;
    int *ts = reinterpret_cast<int *>(melF + a.nPairs * 2 * 64);
    int *lenL = ts + B4;
    int *startL = lenL + B4;                                       // region start; front end: input samples
    int *chanL = startL + ((a.begin || FRONT) ? B4 : 0);           // front end: channels
    long long *offL = reinterpret_cast<long long *>(chanL + (FRONT ? B4 : 0));   // packed: [B][2]
    int *chunkW = reinterpret_cast<int *>(offL + (a.srcOff ? 2 * B4 : 0));   // [16], R2: [32] (upper half at 16)
    float *firW = reinterpret_cast<float *>(chunkW + (R2 ? 32 : 16));   // front end: the 2 lobes FIR taps
    float *dspan = firW + (FRONT ? kFeTaps : 0);                   // front end: [wave][kFeSpan] inputs
    const int segOff = static_cast<int>(dspan + (FRONT ? NW * kFeSpan : 0) - lds);

    {
        const u32x4 *s1 = reinterpret_cast<const u32x4 *>(a.a1), *s2 = reinterpret_cast<const u32x4 *>(a.a2);
        for (int i = tid; i < 256; i += NT) A1[i] = s1[i];
        for (int i = tid; i < 512; i += NT) A2[i] = s2[i];
        for (int i = tid; i < (R2 ? 2 : 1) * kN2 * kN1; i += NT) winP[i] = a.winP[i];
        for (int i = tid; i < 2 * 256; i += NT) tw[i] = a.tw[i];
        const u32x4 *src = reinterpret_cast<const u32x4 *>(a.melFrag);
        for (int i = tid; i < a.nPairs * 2 * 64; i += NT) melF[i] = src[i];
        if (R2) {
            for (int i = tid; i < 2 * 256; i += NT) tw[2 * 256 + i] = a.tw2[i];
            for (int i = tid; i < 16 * NMT; i += NT) w256[i] = i < a.M ? a.w256[i] : 0.0f;
        }
    }
    // per-clip metadata into LDS (the item loop then reads no scalar memory: scalar loads and LDS
    // reads share one completion counter, so a scalar load's latency would stall every LDS wait)
    for (int b = tid; b < a.B; b += NT) {
        if (FRONT) {
            // the resample launch's output length: ceil(samples * outRate / inRate) in FP32
            const int C = max(a.fDims[2 * b + 1], 1);
            const int Lin = max(min(a.fDims[2 * b], static_cast<int>(a.fStride / C)), 0);
            const float lf = static_cast<float>(Lin) * a.fOr;
            lenL[b] = static_cast<int>(ceilf(lf / a.fIr));
            startL[b] = Lin;
            chanL[b] = C;
        } else if (a.srcOff) {
            lenL[b] = max(a.lengths[b], 0);
            offL[2 * b] = a.srcOff[b];
            offL[2 * b + 1] = a.dstOff[b];
        } else {
            const int start = a.begin ? min(max(a.begin[b], 0), static_cast<int>(a.srcStride)) : 0;
            lenL[b] = max(min(a.lengths[b], static_cast<int>(a.srcStride) - start), 0);
            if (a.begin) startL[b] = start;
        }
    }
    if (tid < (R2 ? 32 : 16)) chunkW[tid] = a.pairOff[tid];
    if (FRONT && tid < kFeTaps) {
        // an unclipped output's window values at offsets -lobes .. lobes - 1 (all four-tap groups:
        // the truncating lookup), as the resample launch computes them
        const Win win{a.fLut, a.fWScale, a.fWCenter, a.fLookup};
        firW[tid] = win.trunc_lookup(static_cast<float>(tid - kFeTaps / 2));
    }
    __shared__ int workCtrS;
    int *workCtr = &workCtrS;
    if (tid == 0) workCtrS = 0;
    __syncthreads();
    // the tile plan: items per clip and their exclusive prefix sum in ts[0 .. B]
    {
        __shared__ int wsum[NW];
        int carry = 0;
        for (int base = 0; base < a.B; base += NT) {
            const int b = base + tid;
            const int n = b < a.B ? items_of(a, lenL[b], FI) : 0;
            int v = n;
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(v, o);
                if (lane >= o) v += u;
            }
            if (lane == 63) wsum[wave] = v;
            __syncthreads();
            int before = 0, total = 0;
            for (int w = 0; w < NW; w++) {
                if (w < wave) before += wsum[w];
                total += wsum[w];
            }
            if (b < a.B) ts[b] = carry + before + v - n;
            carry += total;
            __syncthreads();
        }
        if (tid == 0) ts[a.B] = carry;
    }
    __syncthreads();
    const int T = ts[a.B];
    const int off0 = a.center ? a.W / 2 : 0;
    const int nP = a.nPairs;
    float4 px[ROUNDS][4];
    float pp[ROUNDS][4];
    // an item's description (clip b owns item itemIdx)
    auto describe = [&](int itemIdx, int b) {
        Item t;
        t.v = clip_of(a, b, lenL, startL, offL);
        t.f0 = (itemIdx - ts[b]) * FI;
        t.s0 = t.f0 * HOP - off0;
        const int nfr = min(FI, t.v.F - t.f0);
        t.valid = (nfr - 1) * HOP + a.W;
        t.interior = t.s0 >= 1 && t.s0 + t.valid <= t.v.L && (t.valid & 3) == 0;
        return t;
    };
    // the first of the four consecutive samples lane l stages in load m of round r: segment position
    // q = 64 r + 4 (l / 4) + m, phases 4 (l % 4) .. + 3 (nfft 512: samples 16 q + 4 (l % 4) + e); nfft
    // 1024: q = 32 r + 4 (l % 8) + m, phases 4 (l / 8) .. + 3 (samples 32 q + 4 (l / 8) + e), so the
    // eight lanes of one LDS store cycle write one phase row at eight positions
    auto sample_of = [](int r, int m, int lane) {
        return R2 ? 1024 * r + 128 * (lane & 7) + 32 * m + 4 * (lane >> 3)
                  : 1024 * r + 64 * (lane >> 2) + 16 * m + 4 * (lane & 3);
    };
    // an interior item's samples of staging rounds [r0, r1) requested into px/pp (sample_of), and pp
    // the sample before each group; nothing past valid is read.  Interior items only: every group
    // index is clamped to the last valid group, so each load is in bounds and unconditional (no
    // per-lane branch).  A group past valid repeats the last valid group, which is staged too: the
    // item's maximum, and so its scale, is unchanged, and every position such a group fills is one
    // the window zeroes for each frame the item stores.  Buffer loads from the item's sample s0 - 1
    // on: a 32-bit byte offset per lane, and 16-byte loads need only 4-byte alignment.
    auto prefetch = [&](const Item &t, int r0, int r1) {
        const int lane = opaque(static_cast<int>(threadIdx.x)) & 63;
        // (the item's values are wave-uniform; read from lane 0 so the descriptor sits in scalar registers)
        const uintptr_t xa = reinterpret_cast<uintptr_t>(t.v.x + t.s0 - 1);
        const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(xa));
        const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(xa >> 32));
        const int valid = __builtin_amdgcn_readfirstlane(t.valid);
        float *xs = reinterpret_cast<float *>(static_cast<uintptr_t>(lo) | (static_cast<uintptr_t>(hi) << 32));
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(xs, 0, 4 * (valid + 1), 0x00020000);
#pragma unroll
        for (int r = 0; r < ROUNDS; r++) {
            if (r < r0 || r >= r1) continue;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const unsigned o = 4u * static_cast<unsigned>(min(sample_of(r, m, lane), valid - 4));
                px[r][m] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, o + 4u, 0, 0));
                pp[r][m] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, o, 0, 0));
            }
        }
    };
    // Front end: segment slot of item position i (i = sample - s0): the polyphase position
    // (i % 16) PITCH + i / 16; the same formula with an arithmetic shift puts the predecessors
    // i = -1, -2 in the last (spare) column of phase rows 14 and 13.  Clamped, so a lane whose value
    // is discarded never reads outside the segment; branch-free.
    auto slot = [&](int i) {
        i = min(max(i, -2), 16 * (SEGN / 16) - 1);
        return (i & 15) * PITCH + (i >> 4);
    };
    // Front end: the resampled samples [lo, hi) item t reads -- from two before its segment start to
    // its last valid sample (and at least to the reflection of its start), or the whole clip when it
    // is at most half a window long (every index the padding rules reach lies in that range, checked
    // over all clip lengths to 2000 and a sample beyond) -- 128 outputs per round, two per lane, from the
    // down-mixed inputs of the round staged in the wave's span (the next round's loads in flight
    // during this round's products).  Each output is the resample launch's arithmetic: the plain FIR
    // in the HOST order where all 32 taps lie in the clip, the clipped per-tap form at its ends.  The
    // outputs go raw into the segment slots (a short clip: linearly, then copied to the span).
    auto produce = [&](const Item &t, float *seg, int lane) {
        const int b = t.v.b;
        const int Lin = __builtin_amdgcn_readfirstlane(startL[b]);
        const int C = __builtin_amdgcn_readfirstlane(chanL[b]);
        const float *xin = a.fsrc + static_cast<size_t>(b) * a.fStride;
        const float dw = down_mix_weight(C, a.fNormalize);
        constexpr int k = FRONT > 0 ? FRONT : 1;
        const int Lr = t.v.L;
        const bool tiny = !t.interior && Lr <= a.W / 2 + 1;
        const int lo = tiny ? 0 : max(t.s0 - 2, 0);
        const int hi = tiny ? Lr : min(max(t.s0 + t.valid, 1 - t.s0), Lr);   // reflections reach -s0
        float *dsp = dspan + (lane >= 0 ? wave : 0) * kFeSpan;
        float w[kFeTaps];
#pragma unroll
        for (int q = 0; q < kFeTaps; q++) w[q] = firW[q];
        const Win win{a.fLut, a.fWScale, a.fWCenter, a.fLookup};
        const bool pairs = C == 2 && (reinterpret_cast<uintptr_t>(xin) & 7) == 0;
        constexpr int U = kFeSpan / 64;
        float2 v[U];
        auto load = [&](int P0) {
            const int d0 = k * P0 - kFeTaps / 2;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int ti = min(max(d0 + lane + 64 * u, 0), Lin - 1);
                v[u] = *reinterpret_cast<const float2 *>(xin + 2 * static_cast<size_t>(ti));
            }
        };
        if (pairs && lo < hi) load(lo);
        for (int P0 = lo; P0 < hi; P0 += kFeOut) {
            const int d0 = k * P0 - kFeTaps / 2;
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int ti = d0 + lane + 64 * u;
                float d;
                if (pairs) {
                    const float pr[2] = {v[u].x, v[u].y};
                    d = down_mix_sample(pr, 2, dw);
                } else {
                    const int tc = min(max(ti, 0), Lin - 1);
                    d = down_mix_sample(xin + static_cast<size_t>(tc) * C, C, dw);
                }
                dsp[lane + 64 * u] = (ti >= 0 && ti < Lin) ? d : 0.0f;   // outside the clip: never read
            }
            if (pairs && P0 + kFeOut < hi) load(P0 + kFeOut);
            // the lane's two outputs read the k + 32 inputs from 2 k lane on: 8-byte reads into
            // registers, shared by both FIRs
            constexpr int NX = (kFeTaps + k + 1) / 2 * 2;
            float xs[NX];
            {
                const float2 *xp = reinterpret_cast<const float2 *>(dsp + 2 * k * lane);
#pragma unroll
                for (int q = 0; q < NX / 2; q++) {
                    const float2 u = xp[q];
                    xs[2 * q] = u.x;
                    xs[2 * q + 1] = u.y;
                }
            }
#pragma unroll
            for (int uo = 0; uo < 2; uo++) {
                const int P = P0 + 2 * lane + uo;
                const int c = k * P;
                float r;
                if (c - kFeTaps / 2 >= 0 && c + kFeTaps / 2 <= Lin)
                    r = fir_one<kFeTaps>(xs + k * uo, w, 0);
                else
                    r = resample_one([&](int li) { return dsp[li - d0]; }, win, static_cast<float>(c),
                                     max(c - kFeTaps / 2, 0), min(c + kFeTaps / 2, Lin));
                if (P < hi) {
                    if (tiny) seg[P] = r;
                    else seg[slot(P - t.s0)] = r;
                }
            }
        }
        if (tiny)
            for (int i = lane; i < Lr; i += 64) dsp[i] = seg[i];
    };
    // Stages item t into this wave's segment (pre-emphasised, scaled, polyphase) and returns the scale.
    // Interior items from px/pp; items at a clip edge gather each sample with the padding rule (the
    // index and its predecessor's resolved first, both clamped into the clip, so every lane's loads
    // are in flight together).
    auto stage = [&](const Item &t) {
        const int lane = opaque(static_cast<int>(threadIdx.x)) & 63;
        float *seg = lds + segOff + static_cast<size_t>(lane >= 0 ? wave : 0) * SEGF;
        if (FRONT) produce(t, seg, lane);
        float y[ROUNDS][4][4];
        float mx = 0.0f;
        const float c = a.coeff;
        float scale;
        if (t.interior) {
            if (!FRONT && PFR < ROUNDS) prefetch(t, PFR, ROUNDS);   // the rounds not requested ahead
            if (FRONT) {
                // the values prefetch() would have loaded, from the slots
#pragma unroll
                for (int r = 0; r < ROUNDS; r++)
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        const int i = sample_of(r, m, lane);
                        const bool in = i < t.valid;
                        const int ic = min(i, t.valid - 4);
                        const float4 q = make_float4(seg[slot(ic)], seg[slot(ic + 1)], seg[slot(ic + 2)], seg[slot(ic + 3)]);
                        const float pv = seg[slot(ic - 1)];
                        px[r][m] = in ? q : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        pp[r][m] = in ? pv : 0.0f;
                    }
            }
            // the scale from the raw samples: |x[n] - c x[n-1]| <= (1 + |c|) max |x|, so the
            // pre-emphasised samples are formed in the scaled domain at once, one FMA each
            {
                // three running maxima, two new magnitudes per v_max3 each
                float m1 = 0.0f, m2 = 0.0f;
#pragma unroll
                for (int r = 0; r < ROUNDS; r++)
#pragma unroll
                    for (int m = 0; m < 4; m++) {
                        const float4 q = px[r][m];
                        max3_abs(mx, q.x, q.y);
                        max3_abs(m1, q.z, q.w);
                        if (m & 1) max3_abs(m2, pp[r][m - 1], pp[r][m]);
                    }
                mx = fmaxf(mx, fmaxf(m1, m2));
            }
            mx = wave_max_nonneg(mx) * (1.0f + fabsf(c));
            scale = mx > 0.0f ? ldexpf(1.0f, min(9 - ilogbf(mx), 114)) : 1.0f;
#pragma unroll
            for (int r = 0; r < ROUNDS; r++) {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const float4 q = px[r][m];   // zeros past valid stay zero
                    const float x0 = q.x * scale, x1 = q.y * scale, x2 = q.z * scale, x3 = q.w * scale;
                    y[r][m][0] = fmaf(-c, pp[r][m] * scale, x0);
                    y[r][m][1] = fmaf(-c, x0, x1);
                    y[r][m][2] = fmaf(-c, x1, x2);
                    y[r][m][3] = fmaf(-c, x2, x3);
                }
            }
        } else if (t.v.L > a.W / 2 + 1) {
            // a clip edge: one reflection maps every sample a frame reads back into the clip (the
            // frames reach at most W/2 past either end); each index and its predecessor's are resolved
            // without branches and clamped into the clip, so every lane's loads are in flight together
            const Clip &cv = t.v;
            const int L = cv.L;
            const bool refl = a.reflect != 0;
            const int p0 = (a.border == CLAMP) ? 0 : 1;                              // predecessor of sample 0
            const bool p0zero = !(a.border == CLAMP || a.border == REFLECT);
#pragma unroll
            for (int r = 0; r < ROUNDS; r++) {
                __builtin_amdgcn_sched_barrier(0);   // one round's 32 loads in flight at a time
#pragma unroll
                for (int m = 0; m < 4; m++) {
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int i = sample_of(r, m, lane) + e;
                        const int j0 = t.s0 + i;
                        const bool out = j0 < 0 || j0 >= L;
                        int j = j0 < 0 ? -j0 : (j0 >= L ? 2 * (L - 1) - j0 : j0);
                        j = min(max(j, 0), L - 1);
                        const bool zero = i >= t.valid || (out && !refl);
                        const int pj = j == 0 ? p0 : j - 1;
                        const bool pz = j == 0 && p0zero;
                        // both loads unconditional (in-bounds indices); the rules applied by selects
                        const float xv = FRONT ? seg[slot(j - t.s0)] : cv.x[j];
                        const float pl = FRONT ? seg[slot(min(pj, L - 1) - t.s0)] : cv.x[min(pj, L - 1)];
                        const float u = c * (pz ? 0.0f : pl);
                        const float yv = xv - u;
                        y[r][m][e] = zero ? 0.0f : yv;
                        mx = fmaxf(mx, fabsf(y[r][m][e]));
                    }
                }
            }
            mx = wave_max_nonneg(mx);
            // |y * scale| < 2^10 (exponent capped so the scaled domain stays finite for near-silent items)
            scale = mx > 0.0f ? ldexpf(1.0f, min(9 - ilogbf(mx), 114)) : 1.0f;
#pragma unroll
            for (int r = 0; r < ROUNDS; r++)
#pragma unroll
                for (int m = 0; m < 4; m++)
#pragma unroll
                    for (int e = 0; e < 4; e++) y[r][m][e] *= scale;
        } else {
            // a clip no longer than half a window (or empty): sample by sample with RPP's reflect map
            // (any number of bounces), straight into LDS, rescaled below
            const float *xs = FRONT ? dspan + wave * kFeSpan : t.v.x;
            for (int i = lane; i < NPH * PITCH; i += 64) {
                const float yv = (i < t.valid && t.v.L > 0) ? pe_at(xs, t.v.L, t.s0 + i, c, a.border, a.reflect) : 0.0f;
                seg[row_at(i % NPH) + i / NPH] = yv;
                mx = fmaxf(mx, fabsf(yv));
            }
            mx = wave_max_nonneg(mx);
            const float sc = mx > 0.0f ? ldexpf(1.0f, min(9 - ilogbf(mx), 114)) : 1.0f;
            for (int i = lane; i < NPH * PITCH; i += 64) seg[row_at(i % NPH) + i / NPH] *= sc;
            return sc;
        }
#pragma unroll
        for (int r = 0; r < ROUNDS; r++) {
            const int q0 = R2 ? 32 * r + 4 * (lane & 7) : 64 * r + 4 * (lane >> 2);
            if (q0 + 3 < PITCH) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const float4 o = make_float4(y[r][0][e], y[r][1][e], y[r][2][e], y[r][3][e]);
                    const int row = R2 ? 4 * (lane >> 3) + e : 4 * (lane & 3) + e;
                    *reinterpret_cast<float4 *>(&seg[row_at(row) + q0]) = o;
                }
            }
        }
        return scale;
    };
    // the owner of an item: the largest b >= from with ts[b] <= itemIdx
    auto owner = [&](int itemIdx, int from) {
        int lo = from, hi = a.B;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (ts[mid] <= itemIdx) lo = mid;
            else hi = mid;
        }
        return __builtin_amdgcn_readfirstlane(lo);
    };

    // each item is described, and an interior item's samples requested, during its predecessor's mel
    // phase; it is staged at the top of its own iteration
    // Work distribution: workgroup b owns slots j = b, b + G, b + 2G, ... (G workgroups).  Its waves
    // take the first NW in wave order and every later one from a counter in LDS, requested one item
    // ahead, so a wave that runs fast (an easier SIMD partner, fewer clip-edge items) takes more of
    // them.  Workgroups go to the eight XCDs in turn (b % 8), so with G a multiple of 8 slot j runs on
    // XCD j % 8; the slots of one XCD take a contiguous run of items, j / 8 into it: the items running
    // at once on an XCD are neighbours (a clip's consecutive frames), which share their window halo
    // and the 128-byte lines of their output rows in that XCD's L2.  (A workgroup's items ascend
    // either way, which the clip search relies on.)
    const int G = static_cast<int>(gridDim.x);
    const int nMine = T > static_cast<int>(blockIdx.x) ? (T - static_cast<int>(blockIdx.x) + G - 1) / G : 0;
    const bool xcdRuns = (G & 7) == 0;
    auto item_of = [&](int j) {
        if (!xcdRuns) return j;
        const int q = T >> 3, r = T & 7, x = j & 7;
        return x * q + min(x, r) + (j >> 3);
    };
    int k = wave;   // this wave's item, as an index into the workgroup's items
    int item = item_of(static_cast<int>(blockIdx.x) + k * G);
    Item cur{};
    if (k < nMine) {
        cur = describe(item, owner(item, 0));
        if (!FRONT && cur.interior) prefetch(cur, 0, PFR);
    }

    while (k < nMine) {
        int fetched = 0;
        if (lane == 0) fetched = atomicAdd(workCtr, 1);
        const Item it = cur;
        const Clip &v = it.v;
        const int f0 = it.f0;
        const float scale = stage(it);

        // ---- stage 1: per n2, frame fr's samples n1 = 8g .. 8g+7 times the window, split, 2 row
        //      tiles x 3 products; the next n2's reads are issued before this one's split
        __builtin_amdgcn_sched_barrier(0);
        u32x4 a1h[2], a1l[2];
        {
            const int l1 = opaque(static_cast<int>(threadIdx.x)) & 63;
#pragma unroll
            for (int rt = 0; rt < 2; rt++) {
                a1h[rt] = A1[(rt * 2 + 0) * 64 + l1];
                a1l[rt] = A1[(rt * 2 + 1) * 64 + l1];
            }
        }
        // the lane's column: frame l % 16 (nfft 512), or frame l % 8 and parity (l / 8) % 2 (nfft
        // 1024: the parity's samples are phase rows 2 n2 + p, the window's values w[32 n1 + 2 n2 + p])
        const float *rowBase, *winBase;
        constexpr int RS1 = R2 ? 2 * PITCH + 4 : PITCH;   // floats from row n2 to row n2 + 1 of the column
        {
            const int l1 = opaque(static_cast<int>(threadIdx.x));
            const int w1 = l1 >> 6, g1 = (l1 >> 4) & 3;
            const int f1 = R2 ? l1 & 7 : l1 & 15, p1 = (l1 >> 3) & 1;
            rowBase = lds + segOff + w1 * SEGF + (R2 ? row_at(p1) : 0) + f1 * HQ + 8 * g1;
            winBase = winP + (R2 ? p1 * kN2 * kN1 : 0) + 8 * g1;
        }
        float2 dv[2][4];
        float4 wv[2][2];
        auto load1 = [&](int n2, int sl) {
#pragma unroll
            for (int k = 0; k < 4; k++) dv[sl][k] = *reinterpret_cast<const float2 *>(rowBase + n2 * RS1 + 2 * k);
            wv[sl][0] = *reinterpret_cast<const float4 *>(winBase + n2 * kN1);
            wv[sl][1] = *reinterpret_cast<const float4 *>(winBase + n2 * kN1 + 4);
        };
        u32x4 sh[2], slo[2];
        auto split1 = [&](int sl, int o) {
            uint32_t h0, h1, h2, h3, l0, l1, l2, l3;
            split_mul(dv[sl][0].x, wv[sl][0].x, dv[sl][0].y, wv[sl][0].y, h0, l0);
            split_mul(dv[sl][1].x, wv[sl][0].z, dv[sl][1].y, wv[sl][0].w, h1, l1);
            split_mul(dv[sl][2].x, wv[sl][1].x, dv[sl][2].y, wv[sl][1].y, h2, l2);
            split_mul(dv[sl][3].x, wv[sl][1].z, dv[sl][3].y, wv[sl][1].w, h3, l3);
            sh[o] = u32x4{h0, h1, h2, h3};
            slo[o] = u32x4{l0, l1, l2, l3};
        };
        f32x4 acc1[kN2][2];
        load1(0, 0);
        load1(1, 1);
        split1(0, 0);
#pragma unroll
        for (int n2 = 0; n2 < kN2; n2++) {
            // n2 + 2's reads, n2 + 1's split, n2's products
            if (n2 + 2 < kN2) load1(n2 + 2, n2 & 1);
            if (n2 + 1 < kN2) split1((n2 + 1) & 1, (n2 + 1) & 1);
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
                acc1[n2][rt] = mfma3_f16(a1h[rt], a1l[rt], sh[n2 & 1], slo[n2 & 1], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
            __builtin_amdgcn_sched_barrier(0);
        }
        // nfft 1024, bin 512 of a parity's 512-point DFT: sum_n2 (-1)^n2 Y[n2][0] (row 0 of row tile 0:
        // lane group 0), real; bin 256 of the frame is E^2 + O^2 of the two parities' values, then
        // broadcast to every lane of the frame's columns
        float p256 = 0.0f;
        if (R2) {
            float s = 0.0f;
#pragma unroll
            for (int n2 = 0; n2 < kN2; n2++) s += (n2 & 1) ? -acc1[n2][0][0] : acc1[n2][0][0];
            const float so = swap8(s);
            p256 = __shfl(fmaf(s, s, so * so), static_cast<int>(threadIdx.x) & 15);
        }

        // the next item (its request went out at the top of this one) and its clip, a scalar search
        // in the plan (requests are granted in increasing order, so it is not before this clip)
        const int kNext = NW + __builtin_amdgcn_readfirstlane(fetched);
        const int nextItem = kNext < nMine ? item_of(static_cast<int>(blockIdx.x) + kNext * G) : T;
        const int clipNext = nextItem < T ? owner(nextItem, v.b) : 0;

        // ---- twiddle, split, transpose and stage 2, one slot (rt, pair) at a time: the slot's
        //      k1 in lane group g is 8 rt + 2 g + pair; after the transpose piece p is the stage-2
        //      operand of k1 = 8 rt + 2 p + pair
        u32x4 a2h[2][2], a2l[2][2];
        const int l2 = opaque(static_cast<int>(threadIdx.x)) & 63;
        const int g2 = l2 >> 4;
#pragma unroll
        for (int w = 0; w < 2; w++)
#pragma unroll
            for (int o = 0; o < 2; o++) {
                a2h[w][o] = A2[((w * 2 + o) * 2 + 0) * 64 + l2];
                a2l[w][o] = A2[((w * 2 + o) * 2 + 1) * 64 + l2];
            }
        float P[64];
#pragma unroll
        for (int rt = 0; rt < 2; rt++) {
#pragma unroll
            for (int pair = 0; pair < 2; pair++) {
                const int k1 = 8 * rt + 2 * g2 + pair;
                const float4 *twk = reinterpret_cast<const float4 *>(&tw[2 * 16 * k1]);
                float4 t8[8];
#pragma unroll
                for (int u = 0; u < 8; u++) t8[u] = twk[u];   // (c, s) of n2 = 2u and 2u + 1
                uint32_t pcH[4][4], pcL[4][4];   // [piece][register]: piece 0 re n2 0-7, 1 re 8-15, 2 im 0-7, 3 im 8-15
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const float4 t = t8[u];
                    const float re0 = acc1[2 * u][rt][2 * pair], im0 = acc1[2 * u][rt][2 * pair + 1];
                    const float re1 = acc1[2 * u + 1][rt][2 * pair], im1 = acc1[2 * u + 1][rt][2 * pair + 1];
                    const float tr0 = fmaf(re0, t.x, im0 * t.y), ti0 = fmaf(im0, t.x, -(re0 * t.y));
                    const float tr1 = fmaf(re1, t.z, im1 * t.w), ti1 = fmaf(im1, t.z, -(re1 * t.w));
                    split_pair(tr0, tr1, pcH[u >> 2][u & 3], pcL[u >> 2][u & 3]);
                    split_pair(ti0, ti1, pcH[2 + (u >> 2)][u & 3], pcL[2 + (u >> 2)][u & 3]);
                }
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    transpose4(pcH[0][q], pcH[1][q], pcH[2][q], pcH[3][q]);
                    transpose4(pcL[0][q], pcL[1][q], pcL[2][q], pcL[3][q]);
                }
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    const int kk = 8 * rt + 2 * p + pair;       // this product's k1
                    const int which = kk == 0 ? 1 : 0;
                    const u32x4 bh = {pcH[p][0], pcH[p][1], pcH[p][2], pcH[p][3]};
                    const u32x4 bl = {pcL[p][0], pcL[p][1], pcL[p][2], pcL[p][3]};
                    const f32x4 c0 = mfma3_f16(a2h[which][0], a2l[which][0], bh, bl, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
                    const f32x4 c1 = mfma3_f16(a2h[which][1], a2l[which][1], bh, bl, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
                    float q0, q1, q2, q3;
                    if (R2) {
                        // the frame's other parity holds the same bins in lane l ^ 8: |V z + y|^2 with
                        // z this lane's value, y the other's and V the twiddle of the bin (its real
                        // part negated for parity 1)
                        const float4 t01 = tw2[(g2 * 16 + kk) * 2], t23 = tw2[(g2 * 16 + kk) * 2 + 1];
                        const float sg = (l2 & 8) ? -1.0f : 1.0f;
                        auto rad2 = [&](float zr, float zi, float vr, float vi) {
                            const float yr = swap8(zr), yi = swap8(zi);
                            vr *= sg;
                            const float xr = fmaf(vr, zr, fmaf(-vi, zi, yr));
                            const float xi = fmaf(vr, zi, fmaf(vi, zr, yi));
                            return fmaf(xi, xi, xr * xr);
                        };
                        q0 = rad2(c0[0], c0[1], t01.x, t01.y);
                        q1 = rad2(c0[2], c0[3], t01.z, t01.w);
                        q2 = rad2(c1[0], c1[1], t23.x, t23.y);
                        q3 = rad2(c1[2], c1[3], t23.z, t23.w);
                    } else {
                        q0 = fmaf(c0[1], c0[1], c0[0] * c0[0]);
                        q1 = fmaf(c0[3], c0[3], c0[2] * c0[2]);
                        q2 = fmaf(c1[1], c1[1], c1[0] * c1[0]);
                        q3 = fmaf(c1[3], c1[3], c1[2] * c1[2]);
                    }
                    if (which) {
                        P[0] = q0; P[16] = q1; P[32] = q2; P[48] = q3;
                    } else {
                        P[kk] = q0; P[kk + 32] = q1; P[32 - kk] = q2; P[64 - kk] = q3;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if (GENERAL && a.power == 1) {
#pragma unroll
            for (int o = 0; o < 64; o++) P[o] = sqrtf(P[o]);
        }

        // ---- power -> bf16 hi/lo pairs, transposed so lane group g holds bins 32 c + 8 g .. + 7
        uint32_t PH[32], PL[32];
#pragma unroll
        for (int k = 0; k < 32; k++) split_bf16_pair(P[2 * k], P[2 * k + 1], PH[k], PL[k]);
#pragma unroll
        for (int e = 0; e < 2; e++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                transpose4(PH[16 * e + q], PH[16 * e + 4 + q], PH[16 * e + 8 + q], PH[16 * e + 12 + q]);
                transpose4(PL[16 * e + q], PL[16 * e + 4 + q], PL[16 * e + 8 + q], PL[16 * e + 12 + q]);
            }

        // ---- the next item's samples requested now, so the mel phase, the dB stores and the loop
        //      turn cover their latency (and the next item's load wait does not include this
        //      item's stores, which are issued after them)
        Item nxt{};
        if (nextItem < T) {
            nxt = describe(nextItem, clipNext);
            if (!FRONT && nxt.interior) prefetch(nxt, 0, PFR);
        }

        if constexpr (!GENERAL) {
            // ---- banded mel on MFMA, streamed over the filter tiles, then dB and the stores.  The runs
            //      (a chunk's tiles; nfft 1024: the lower half's runs by ascending chunk, then the upper
            //      half's by descending chunk -- bins 512 - u, on the power of the lane's partner l ^ 8,
            //      so the frame's parity-0 column gathers the whole filter bank) ascend in frequency, and
            //      so does each run's first tile (the table builder checks it): a window of MW
            //      accumulators holds tiles lo .. lo + MW - 1, and a tile below the current run's
            //      first tile is complete, so it is converted to dB (nfft 1024: bin 256's weights added
            //      first) and stored, and the window moves up.  Chunk c's operand is
            //      PH/PL[16 (c % 2) + 4 (c / 2) + 0..3]; a run's first two pairs' weight fragments are
            //      requested one run ahead, into the register set of the run's parity, the others when
            //      used.
            constexpr int NR = R2 ? 2 * kChunks : kChunks;
            constexpr int MW = mel_window(NMT);
            int runW[NR];   // all read at once: one LDS wait, not one per run
#pragma unroll
            for (int r = 0; r < NR; r++) runW[r] = chunkW[r < kChunks ? r : 16 + (NR - 1 - r)];
            const u32x4 *melL = melF + (opaque(static_cast<int>(threadIdx.x)) & 63);
            u32x4 wf[2][2][2];   // [run parity][first, second pair][hi, lo]
            auto request = [&](int w, int par) {
                const int p0 = w >> 16;
                const int q0 = max(min(p0, nP - 1), 0), q1 = max(min(p0 + 1, nP - 1), 0);
                wf[par][0][0] = melL[(q0 * 2) * 64];
                wf[par][0][1] = melL[(q0 * 2 + 1) * 64];
                wf[par][1][0] = melL[(q1 * 2) * 64];
                wf[par][1][1] = melL[(q1 * 2 + 1) * 64];
            };
            request(__builtin_amdgcn_readfirstlane(runW[0]), 0);
            // lane l stores frame f0 + l % 16 (nfft 1024: f0 + l % 8, from the parity-0 lanes), rows
            // 16 t + 4 g + i of tile t: a scalar row base plus the lane's offset 4 g hs + f
            const int lane = opaque(static_cast<int>(threadIdx.x)) & 63;
            const int fr = R2 ? lane & 7 : lane & 15, g = lane >> 4;
            const int f = f0 + fr;
            const bool storeLane = f < v.F && !(R2 && (lane & 8));
            const uint32_t laneOff = 4u * static_cast<uint32_t>(g) * v.hs + static_cast<uint32_t>(f);
            const bool off32 = v.hs < (1u << 26);   // the lane's byte offset fits 32 bits
            // the planes hold power of operands scaled by scale * 2^12 (scale a power of two): undo once
            // per output
            const float unscale = ldexpf(1.0f, -12 - ilogbf(scale));
            const float kdb = unscale * unscale * a.invRef;
            const int nMT = (a.M + 15) >> 4;
            f32x4 acc[MW];
#pragma unroll
            for (int j = 0; j < MW; j++) acc[j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            int lo = 0;
            auto retire = [&]() {
                f32x4 t = acc[0];
                if (R2) {
                    const float4 w = *reinterpret_cast<const float4 *>(w256 + 16 * lo + 4 * g);
                    t[0] = fmaf(w.x, p256, t[0]);
                    t[1] = fmaf(w.y, p256, t[1]);
                    t[2] = fmaf(w.z, p256, t[2]);
                    t[3] = fmaf(w.w, p256, t[3]);
                }
                if (storeLane) {
                    if (off32 && 16 * lo + 16 <= a.M) {
                        const uint32_t laneB = 4u * laneOff;
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const float val = a.mult * __builtin_amdgcn_logf(fmaxf(a.minRatio, t[i] * kdb));
                            char *rowp = reinterpret_cast<char *>(v.out + static_cast<size_t>(16 * lo + i) * v.hs);
                            *reinterpret_cast<float *>(rowp + laneB) = val;
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const float val = a.mult * __builtin_amdgcn_logf(fmaxf(a.minRatio, t[i] * kdb));
                            float *rowp = v.out + static_cast<size_t>(16 * lo + i) * v.hs;
                            if (16 * lo + 4 * g + i < a.M) rowp[laneOff] = val;
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j + 1 < MW; j++) acc[j] = acc[j + 1];
                acc[MW - 1] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                lo++;
            };
#pragma unroll
            for (int r = 0; r < NR; r++) {
                const int w = __builtin_amdgcn_readfirstlane(runW[r]);
                if (r + 1 < NR) request(__builtin_amdgcn_readfirstlane(runW[r + 1]), (r + 1) & 1);
                const int t0 = w & 255, t1 = (w >> 8) & 255, p0 = w >> 16;
                if (t0 >= t1) continue;
                while (lo < t0) retire();
                const int c = r < kChunks ? r : NR - 1 - r;
                const int base = 16 * (c & 1) + 4 * (c >> 1);
                u32x4 bh = {PH[base], PH[base + 1], PH[base + 2], PH[base + 3]};
                u32x4 bl = {PL[base], PL[base + 1], PL[base + 2], PL[base + 3]};
                if (R2 && r >= kChunks) {
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        bh[e] = swap8u(bh[e]);
                        bl[e] = swap8u(bl[e]);
                    }
                }
                acc[0] = mfma3_bf16(wf[r & 1][0][0], wf[r & 1][0][1], bh, bl, acc[0]);
                if (t1 - t0 > 1) acc[1] = mfma3_bf16(wf[r & 1][1][0], wf[r & 1][1][1], bh, bl, acc[1]);
#pragma unroll
                for (int j = 2; j < MW; j++)
                    if (t1 - t0 > j) {
                        const int q = p0 + j;
                        acc[j] = mfma3_bf16(melL[(q * 2) * 64], melL[(q * 2 + 1) * 64], bh, bl, acc[j]);
                    }
            }
            while (lo < nMT) retire();
        } else {
            // ---- banded mel on MFMA, pairs in their table order (chunk by chunk, tiles ascending): chunk
            //      c's operand is PH/PL[16 (c % 2) + 4 (c / 2) + 0..3]; each pair's weight fragments were
            //      requested two pairs ahead
            //      (nfft 1024: each chunk's second run is the upper half's, on the power of the lane's
            //      partner l ^ 8 -- the frame's bins 512 - u -- so the frame's parity-0 column gathers the
            //      whole filter bank; then bin 256's weights)
            f32x4 mel[NMT];
#pragma unroll
            for (int mt = 0; mt < NMT; mt++) mel[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            const u32x4 *melL = melF + (opaque(static_cast<int>(threadIdx.x)) & 63);
            u32x4 wch = melL[0], wcl = melL[64];
            u32x4 wnh = melL[(min(1, nP - 1) * 2) * 64], wnl = melL[(min(1, nP - 1) * 2 + 1) * 64];
            int pn = 2;
            auto run = [&](int t0, int t1, const u32x4 &bh, const u32x4 &bl) {
#pragma unroll
                for (int mt = 0; mt < NMT; mt++) {
                    if (mt >= t0 && mt < t1) {
                        mel[mt] = mfma3_bf16(wch, wcl, bh, bl, mel[mt]);
                        wch = wnh;
                        wcl = wnl;
                        const int q = min(pn, nP - 1);
                        wnh = melL[(q * 2) * 64];
                        wnl = melL[(q * 2 + 1) * 64];
                        pn++;
                    }
                }
            };
#pragma unroll
            for (int ch = 0; ch < kChunks; ch++) {
                const int cw = __builtin_amdgcn_readfirstlane(chunkW[ch]);   // read per item from LDS
                const int base = 16 * (ch & 1) + 4 * (ch >> 1);
                const u32x4 bh = {PH[base], PH[base + 1], PH[base + 2], PH[base + 3]};
                const u32x4 bl = {PL[base], PL[base + 1], PL[base + 2], PL[base + 3]};
                run(cw & 255, (cw >> 8) & 255, bh, bl);
                if (R2) {
                    const int cw2 = __builtin_amdgcn_readfirstlane(chunkW[16 + ch]);
                    const int t0 = cw2 & 255, t1 = (cw2 >> 8) & 255;
                    if (t0 < t1) {   // (no partner swap for a chunk without pairs)
                        u32x4 sh, sl;
#pragma unroll
                        for (int e = 0; e < 4; e++) {
                            sh[e] = swap8u(bh[e]);
                            sl[e] = swap8u(bl[e]);
                        }
                        run(t0, t1, sh, sl);
                    }
                }
            }
            if (R2) {
                const int g = (opaque(static_cast<int>(threadIdx.x)) >> 4) & 3;
#pragma unroll
                for (int mt = 0; mt < NMT; mt++) {
                    const float4 w = *reinterpret_cast<const float4 *>(w256 + 16 * mt + 4 * g);
                    mel[mt][0] = fmaf(w.x, p256, mel[mt][0]);
                    mel[mt][1] = fmaf(w.y, p256, mel[mt][1]);
                    mel[mt][2] = fmaf(w.z, p256, mel[mt][2]);
                    mel[mt][3] = fmaf(w.w, p256, mel[mt][3]);
                }
            }


            // ---- dB (or the power, for the clip-maximum reference) and the stores
            // the planes hold power of operands scaled by scale * 2^12: undo once per output
            const float unscale = ldexpf(1.0f, -12 - ilogbf(scale));   // scale is a power of two
            const float melScale = (GENERAL && a.power == 1) ? unscale : unscale * unscale;
            const int lane = opaque(static_cast<int>(threadIdx.x)) & 63;
            const int fr = R2 ? lane & 7 : lane & 15, g = lane >> 4;
            const int f = f0 + fr;
            float cmax = -3.40282347e38f;
#pragma unroll
            for (int mt = 0; mt < NMT; mt++) {
                const int m0 = 16 * mt + 4 * g;
                float val[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float pw = mel[mt][i] * melScale;
                    if (a.clipMax) {
                        val[i] = pw;
                        if (f < v.F && m0 + i < a.M) cmax = fmaxf(cmax, pw);
                    } else if (a.minRatio >= 1.17549435e-38f) {
                        val[i] = a.mult * __builtin_amdgcn_logf(fmaxf(a.minRatio, pw * a.invRef));
                    } else {
                        const float r = pw * a.invRef;
                        val[i] = a.mult * log2f(a.minRatio < r ? r : a.minRatio);
                    }
                }
                if (f < v.F) {
                    if (a.ntf) {
                        float *o = v.out + static_cast<size_t>(f) * v.hs + m0;
#pragma unroll
                        for (int i = 0; i < 4; i++)
                            if (m0 + i < a.M) o[i] = val[i];
                    } else {
#pragma unroll
                        for (int i = 0; i < 4; i++)
                            if (m0 + i < a.M) v.out[static_cast<size_t>(m0 + i) * v.hs + f] = val[i];
                    }
                }
            }
            if (a.clipMax) {
                cmax = wave_max(cmax);
                if (lane == 0) atomicMax(&a.clipMax[v.b], __float_as_int(cmax));
            }
        }

        cur = nxt;
        k = kNext;
    }
}

}  // namespace

// ------------------------------------------------------------------------- host side
// a1 4 KB, a2 8 KB, winP 2 KB (nfft 1024: 4 KB), tw 2 KB, nfft 1024: tw2 2 KB, then W floats for the
// window itself
size_t fft_table_bytes(int W, int nfft)
{
    return 4096 + 8192 + (nfft == 2 * kN ? 4096 + 2048 : 2048) + 2048 + sizeof(float) * static_cast<size_t>(W);
}

void fft_tables_carve(char *mem, int W, FftTables &t, int nfft)
{
    const bool r2 = nfft == 2 * kN;
    t.a1 = reinterpret_cast<uint16_t *>(mem);
    t.a2 = reinterpret_cast<uint16_t *>(mem + 4096);
    t.winP = reinterpret_cast<float *>(mem + 4096 + 8192);
    t.tw = t.winP + (r2 ? 2 : 1) * kN2 * kN1;
    t.tw2 = r2 ? t.tw + 512 : nullptr;
    t.mem = mem;
    t.nfft = r2 ? 2 * kN : kN;
    t.W = W;
}

hipError_t build_fft_tables_at(hipStream_t st, int W, FftTables &t)
{
    float *win = (t.tw2 ? t.tw2 : t.tw) + 512;
    hipLaunchKernelGGL(fft_hann_kernel, dim3(ceil_div(W, 256)), dim3(256), 0, st, W, win);
    hipLaunchKernelGGL(fft_tables_kernel, dim3(8), dim3(256), 0, st, W, t.nfft, win, t.a1, t.a2, t.winP, t.tw, t.tw2);
    return hipGetLastError();
}

size_t mel_frag_bytes(int M, int nfft)
{
    return static_cast<size_t>((M + 15) / 16) * kChunks * 2048 * (nfft == 2 * kN ? 2 : 1);
}
size_t mel_frag_word_bytes(int nfft) { return sizeof(int) * 257 + (nfft == 2 * kN ? sizeof(float) * 16 * kMaxMT : 0); }

hipError_t build_mel_frag_at(hipStream_t st, const MelTables &mt, MelFrag &f)
{
    hipLaunchKernelGGL(mel_frag_kernel, dim3(1), dim3(256), 0, st, mt.dense, mt.filterBins, mt.M, mt.numBins,
                       f.w256 ? 1 : 0, f.frag, f.pairOff, f.w256);
    return hipGetLastError();
}

const FftTables *get_fft_tables(rppHandle *h, int nfft, int W)
{
    const auto key = std::make_pair(nfft, W);
    auto it = h->fft.find(key);
    if (it != h->fft.end()) return &it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return nullptr;
    char *mem = nullptr;
    if (hipMalloc(&mem, fft_table_bytes(W, nfft)) != hipSuccess) return nullptr;
    FftTables t;
    fft_tables_carve(mem, W, t, nfft);
    if (build_fft_tables_at(h->stream, W, t) != hipSuccess) return nullptr;
    h->fft[key] = t;
    return &h->fft[key];
}

const MelFrag *get_mel_frag(rppHandle *h, const MelTables *mt, const std::string &key)
{
    auto it = h->melFrag.find(key);
    if (it != h->melFrag.end()) return &it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return nullptr;
    MelFrag f;
    const int nfft = mt->numBins == kN + 1 ? 2 * kN : kN;   // nfft 1024 (513 bins): the two halves and bin 256
    const bool r2 = nfft == 2 * kN;
    if (hipMalloc(&f.frag, mel_frag_bytes(mt->M, nfft)) != hipSuccess) return nullptr;
    if (hipMalloc(&f.pairOff, mel_frag_word_bytes(nfft)) != hipSuccess) return nullptr;
    if (r2) f.w256 = reinterpret_cast<float *>(f.pairOff + 257);
    if (build_mel_frag_at(h->stream, *mt, f) != hipSuccess) return nullptr;
    int n[2] = {0, 0};   // [255] tiles of a chunk not contiguous, [256] pair count
    if (hipMemcpyAsync(n, f.pairOff + 255, sizeof(n), hipMemcpyDeviceToHost, h->stream) != hipSuccess) return nullptr;
    if (hipStreamSynchronize(h->stream) != hipSuccess) return nullptr;
    f.nPairs = n[0] ? 0 : n[1];   // 0: the factored kernel cannot use this filter bank
    h->melFrag[key] = f;
    return &h->melFrag[key];
}

// LDS bytes of the factored kernel for nfft, hop H, B clips and nPairs mel pairs; nfft 1024 with
// `waves` waves per workgroup (0: the fewest it runs with).
size_t fft_lds_bytes(int nfft, int H, int B, int nPairs, bool begin, bool packed, int *pitchOut, bool front, int waves)
{
    const size_t B4 = static_cast<size_t>((B + 1 + 3) / 4 * 4);
    const size_t perClip = sizeof(int) * B4 * (2 + ((begin || front) ? 1 : 0) + (front ? 1 : 0)) +
                           (packed ? sizeof(long long) * 2 * B4 : 0);
    if (nfft == 2 * kN) {
        if (pitchOut) *pitchOut = kPitchR2;
        // a1, a2, winP [2][16][32], tw, tw2, bin 256's weights, both halves' fragments, the plan,
        // 32 chunk words, the segments
        return 4096 + 8192 + sizeof(float) * (2 * kN2 * kN1 + 512 + 512 + 16 * kMaxMT) +
               static_cast<size_t>(nPairs) * 2048 + perClip + sizeof(int) * 32 +
               sizeof(float) * (waves > 0 ? waves : kWavesR2Min) * static_cast<size_t>(32 * kPitchR2 + 64);
    }
    const int pitch = pitch_for(((kFI - 1) * H + kN) / 16);
    if (pitchOut) *pitchOut = pitch;
    const size_t frontBytes = front ? sizeof(float) * (kFeTaps + kWaves * kFeSpan) : 0;
    return 4096 + 8192 + sizeof(float) * (kN2 * kN1 + 512) + static_cast<size_t>(nPairs) * 2048 + perClip +
           sizeof(int) * 16 + frontBytes + sizeof(float) * kWaves * 16 * static_cast<size_t>(pitch);
}

// fp16x3 with nfft 512, W <= 512 and hop 160, or nfft 1024, W <= 1024 and hop 256 (the hops the
// kernel is instantiated for), at most 128 filters and 4096 clips (the tile plan lives in LDS)
bool fft_eligible(const RpptLogMelParams *p, RpptLogMelPrecision precision, int B)
{
    const int nfft = p->nfft ? p->nfft : p->windowLength;
    if (precision != RPPT_LOGMEL_FP16X3 || (nfft != kN && nfft != 2 * kN) || p->windowLength > nfft ||
        p->windowLength <= 0)
        return false;
    if (p->windowStep != (nfft == kN ? 160 : 256)) return false;
    return p->numFilter <= 16 * kMaxMT && B <= 4096;
}

bool fft_r2_launchable(const RpptLogMelParams *p, int ntf, bool clipRef, bool frontEnd)
{
    const float minRatio = static_cast<float>(std::pow(10.0, static_cast<double>(p->cutOffDB / p->multiplier)));
    return p->power == 2 && !ntf && !clipRef && !frontEnd && minRatio >= 1.17549435e-38f;
}

bool fft_front_eligible(int k, int lobes) { return k >= 2 && k <= kFeMaxK && 2 * lobes == kFeTaps; }

RppStatus logmel_fft_launch(rppHandle *h, const FftLaunch &L)
{
    const RpptLogMelParams *p = L.p;
    int pitch = 0;
    const bool front = L.fe != nullptr;
    if (front && (L.begin || L.srcOff)) return RPP_ERROR_NOT_IMPLEMENTED;
    const int nfft = L.ft->nfft;
    const bool r2 = nfft == 2 * kN;
    // nfft 1024: eight waves per workgroup where their segments fit beside the mel table and the
    // clip plan, else six
    const auto r2Bytes = [&](int w) {
        return fft_lds_bytes(nfft, p->windowStep, L.B, L.mf->nPairs, L.begin != nullptr, L.srcOff != nullptr, &pitch,
                             front, w);
    };
    const int r2Waves = r2 && r2Bytes(kWavesR2) <= static_cast<size_t>(h->maxLdsOptin) ? kWavesR2 : kWavesR2Min;
    const size_t bytes = r2 ? r2Bytes(r2Waves)
                            : fft_lds_bytes(nfft, p->windowStep, L.B, L.mf->nPairs, L.begin != nullptr,
                                            L.srcOff != nullptr, &pitch, front);
    if (bytes > static_cast<size_t>(h->maxLdsOptin)) return RPP_ERROR_OUT_OF_BOUND_SHARED_MEMORY_SIZE;
    FftArgs a{};
    a.src = L.src;
    a.srcStride = L.srcStride;
    a.lengths = L.lengths;
    a.begin = L.begin;
    a.B = L.B;
    a.srcOff = L.srcOff;
    a.dstOff = L.dstOff;
    a.dst = L.dst;
    a.dstNStride = L.dstNStride;
    a.hStride = L.hStride;
    a.ntf = L.ntf;
    a.frameCap = L.frameCap;
    a.clipMax = L.clipMax;
    a.coeff = p->preEmphasisCoeff;
    a.border = static_cast<int>(p->borderType);
    a.W = p->windowLength;
    a.H = p->windowStep;
    a.center = p->centerWindows;
    a.reflect = p->reflectPadding;
    a.power = p->power == 2 ? 2 : 1;
    a.M = p->numFilter;
    a.minRatio = L.minRatio;
    a.mult = L.mult;
    a.invRef = L.invRef;
    a.a1 = L.ft->a1;
    a.a2 = L.ft->a2;
    a.winP = L.ft->winP;
    a.tw = L.ft->tw;
    a.melFrag = L.mf->frag;
    a.pairOff = L.mf->pairOff;
    a.nPairs = L.mf->nPairs;
    a.tw2 = L.ft->tw2;
    a.w256 = L.mf->w256;
    a.pitch = pitch;
    if (front) {
        a.src = L.fe->src;   // only offset, never read: the kernel reads fsrc
        a.fsrc = L.fe->src;
        a.fStride = L.fe->stride;
        a.fDims = L.fe->dims;
        a.fK = L.fe->k;
        a.fNormalize = L.fe->normalize;
        a.fIr = L.fe->inRate;
        a.fOr = L.fe->outRate;
        a.fLut = L.fe->lut;
        a.fWScale = L.fe->wScale;
        a.fWCenter = L.fe->wCenter;
        a.fLookup = L.fe->lookupSize;
    }
    if (L.mf->nPairs <= 0) return RPP_ERROR_NOT_IMPLEMENTED;
    const int nMT = (p->numFilter + 15) / 16;
    // one workgroup per CU, or one per item when there are fewer items than CUs (each item's wave
    // then has a compute unit to itself: workload 0's 105 items in 16.1 us instead of 25.0 us as 14
    // full workgroups); the host bounds the items by the row stride (strided layout; the packed
    // layout's lengths are on the device)
    unsigned grid = static_cast<unsigned>(h->numCU);
    if (!L.srcOff) {
        long long maxLen = L.srcStride;
        if (front) maxLen = static_cast<long long>(std::ceil(static_cast<double>(L.fe->stride) * L.fe->outRate / L.fe->inRate)) + 1;
        const long long maxF = std::min<long long>(std::max(num_windows(static_cast<int>(std::min<long long>(maxLen, 0x7fffffffLL)),
                                                                        p->windowLength, p->windowStep, p->centerWindows != 0),
                                                            0),
                                                   L.dstOff ? 0x7fffffffLL : static_cast<long long>(L.frameCap));
        const int fi = r2 ? 8 : kFI;
        const long long items = static_cast<long long>(L.B) * ((maxF + fi - 1) / fi);
        grid = static_cast<unsigned>(std::max<long long>(1, std::min<long long>(grid, items)));
    }
    const bool general = a.power != 2 || a.ntf || a.clipMax || !(a.minRatio >= 1.17549435e-38f);
    // the front end runs with the common launch only (the caller falls back to two launches)
    if (front && general) return RPP_ERROR_NOT_IMPLEMENTED;
#define RPPA_FFT_LAUNCH(NMT, GEN, FE)                                                                        \
    do {                                                                                                     \
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&logmel_fft_kernel<NMT, 160, GEN, FE>),     \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));      \
        hipLaunchKernelGGL((logmel_fft_kernel<NMT, 160, GEN, FE>), dim3(grid), dim3(kThreads), bytes, h->stream, a); \
    } while (0)
    const int fk = front ? L.fe->k : 0;
    if (front && fk != 2 && fk != 3) return RPP_ERROR_NOT_IMPLEMENTED;
    if (r2) {
        if (general || front || p->windowStep != 256 || !a.tw2 || !a.w256) return RPP_ERROR_NOT_IMPLEMENTED;
#define RPPA_FFT2_LAUNCH(NMT, NW)                                                                                  \
    do {                                                                                                           \
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&logmel_fft_kernel<NMT, 256, false, 0, 1024, NW>), \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));            \
        hipLaunchKernelGGL((logmel_fft_kernel<NMT, 256, false, 0, 1024, NW>), dim3(grid), dim3(64 * NW), bytes,     \
                           h->stream, a);                                                                          \
    } while (0)
        if (r2Waves == kWavesR2) {
            if (nMT <= 5) RPPA_FFT2_LAUNCH(5, kWavesR2);
            else RPPA_FFT2_LAUNCH(8, kWavesR2);
        } else {
            if (nMT <= 5) RPPA_FFT2_LAUNCH(5, kWavesR2Min);
            else RPPA_FFT2_LAUNCH(8, kWavesR2Min);
        }
#undef RPPA_FFT2_LAUNCH
        RPPA_HIP(hipGetLastError());
        return RPP_SUCCESS;
    }
    if (p->windowStep != 160) return RPP_ERROR_NOT_IMPLEMENTED;
    if (nMT <= 5) {
        if (fk == 3) RPPA_FFT_LAUNCH(5, false, 3);
        else if (fk == 2) RPPA_FFT_LAUNCH(5, false, 2);
        else if (general) RPPA_FFT_LAUNCH(5, true, 0);
        else RPPA_FFT_LAUNCH(5, false, 0);
    } else {
        if (fk == 3) RPPA_FFT_LAUNCH(8, false, 3);
        else if (fk == 2) RPPA_FFT_LAUNCH(8, false, 2);
        else if (general) RPPA_FFT_LAUNCH(8, true, 0);
        else RPPA_FFT_LAUNCH(8, false, 0);
    }
#undef RPPA_FFT_LAUNCH
    RPPA_HIP(hipGetLastError());
    return RPP_SUCCESS;
}

}  // namespace rppa
