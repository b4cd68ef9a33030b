// logmel_fused.hip -- the whole log-mel chain over a ragged batch in one persistent launch.
//
// Equivalent to rppt_pre_emphasis_filter_gpu -> rppt_spectrogram_gpu (NFT, power) ->
// rppt_mel_filter_bank_gpu -> rppt_to_decibels_gpu with matching arguments, without the three
// HBM intermediates: per work item (FT frames of one clip) the kernel
//   1. stages the clip segment those frames cover in LDS, applying pre-emphasis and the
//      reflect/zero padding on the fly (HBM: the raw samples, read once);
//   2. runs the DFT as a contraction of the implicit frame matrix [FT x W] with the
//      window-folded basis [W x N] (tables.hip) -- on MFMA (bf16/fp16, split hi+lo,
//      FP32 accumulation) or on FP32 FMA;
//   3. forms power from the re/im column pairs into an LDS tile [FT x K];
//   4. projects onto the mel filters -- MFMA (bf16 split hi+lo) or the exact FP32 gather;
//   5. converts to dB and stores the NFT output (HBM: the dB values, written once).
//
// Work list: a planner kernel turns the clip lengths into a prefix sum of tiles per clip; the
// persistent grid strides over the tiles.  Everything stays on the device, so the two launches
// can be captured in a hipGraph and replayed for any batch living in the same buffers.
//
// LDS layouts.  Signal segment: sample i of the segment at i + (i / H) * 4 -- a 16-byte skew per
// hop, so the 32 frames read by a half-wave start in different banks (frames are H apart and
// H*4 bytes is a multiple of the bank period).  Power tile: [FT][KP + 4].
#include "common.hpp"

#include <cmath>

namespace rppa {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kSkew = 4;

struct LogMelArgs {
    const float *src;
    uint32_t srcStride;
    const int *lengths;
    const int *begin;      // first sample of each clip's region in its row (rows start at 0 if null)
    int B;
    float *dst;
    uint32_t dstNStride, hStride;
    float coeff;
    int border;
    int N, W, H, center, reflect, power;
    int K, M;
    float minRatio, mult, invRef;
    int ntf;               // output [n][frames][hStride] (else [n][M][hStride]); never with the packed layout
    uint32_t frameCap;     // frames a clip's output holds (hStride for NFT, the descriptor's rows for NTF)
    int *clipMax;          // per-clip maximum of the mel power (float bits) when the reference is the
                           // clip maximum: the kernel then stores power and a second pass converts to dB
    const int *tileStart;  // [B + 1]
    const int4 *tileRec;   // [tiles] clip, first frame, region length, region start of each tile (null:
                           // the MFMA kernel searches tileStart instead)
    int4 *blockRec;        // when set, the MFMA kernel plans in each block (no planner launch) and keeps
    int kPad;              // block b's records at blockRec[b * kPad + k] (kPad a multiple of 16)
    // packed layout (both null for the [n][nStride] / [n][M][hStride] descriptors): clip b's samples
    // at src + srcOff[b], its dB rows at dst + dstOff[b] as a dense [M][frames_b] block
    const long long *srcOff, *dstOff;
    // FP32 basis
    const float *basis;
    int cols;
    // split basis
    const uint16_t *bHi, *bLo;      // folded split basis (DftTables::foldHi/foldLo)
    const uint16_t *bLo2;           // its third term (bf16 three-term split)
    const float *wHalf;
    int Wp;                         // folded contraction length, padded to 16
    float basisInv;                 // 1 / the scale of the fp16 basis (an exact power of two)
    // mel
    const float *weightsDown;
    const int *intervals;
    const float *normFactors;
    const int *filterBins;
    int normalize;
    const uint16_t *mHi, *mLo;
    int KP, Mpad;
};

__host__ __device__ __forceinline__ int seg_pos(int i, int H) { return i + (i / H) * kSkew; }

// The value of v, hidden from the optimiser: thread-index arithmetic recomputed from it inside the
// tile loop stays there instead of being hoisted into registers that live across the whole kernel.
__device__ __forceinline__ int opaque(int v)
{
    __asm__ volatile("" : "+v"(v));
    return v;
}

// Maximum over the wave, with the exchange addresses formed from a fresh opaque lane index (the
// plain shuffles' addresses are loop-invariant and would be hoisted into live registers).
__device__ __forceinline__ float wave_max(float v)
{
    const int lane = opaque(static_cast<int>(threadIdx.x)) & 63;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v = fmaxf(v, __int_as_float(__builtin_amdgcn_ds_bpermute((lane ^ o) << 2, __float_as_int(v))));
    return v;
}

// seg_pos for 0 <= i < 2^20 without an integer division: with i = qH + r, (i + 0.5)/H lies at least
// 0.5/H away from an integer, far beyond the FP32 rounding of the product.
__device__ __forceinline__ int seg_pos_r(int i, float invH)
{
    return i + static_cast<int>((static_cast<float>(i) + 0.5f) * invH) * kSkew;
}

// Where clip b lives: samples, length, frame count, output base and output row stride.
struct ClipView {
    const float *x;
    float *out;
    uint32_t hs;
    int L, F, b;
};

// Element i of an array that stays unchanged while the kernel runs (the per-clip inputs, the tile
// plan).  Read through the constant address space: with a wave-uniform index the compiler issues a
// scalar load, which neither waits for nor holds up the vector loads in flight (a divergent index
// still gets a vector load).
template <typename T>
__device__ __forceinline__ T ld_const(const T *p, int i)
{
    typedef const __attribute__((address_space(4))) T *ConstPtr;
    return ((ConstPtr)p)[i];
}

// Region of clip b's row: samples [start, start + length) with start = begin[b] (0 without a
// begin array) clamped to the row, and the length clamped so the region never leaves the row.
__device__ __forceinline__ int row_region(const int *begin, const int *lengths, int b, uint32_t srcStride, int &start)
{
    start = begin ? min(max(ld_const(begin, b), 0), static_cast<int>(srcStride)) : 0;
    return max(min(ld_const(lengths, b), static_cast<int>(srcStride) - start), 0);
}

// Frame count and output placement of clip b, whose samples (length v.L) are already set in v.
__device__ __forceinline__ ClipView clip_view_out(const LogMelArgs &a, ClipView v, int b)
{
    int F = max(num_windows(v.L, a.W, a.H, a.center != 0), 0);
    if (a.dstOff) {
        v.out = a.dst + ld_const(a.dstOff, b);
        v.hs = static_cast<uint32_t>(F);
    } else {
        v.out = a.dst + static_cast<size_t>(b) * a.dstNStride;
        v.hs = a.hStride;
        F = min(F, static_cast<int>(a.frameCap));
    }
    v.F = F;
    v.b = b;
    return v;
}

__device__ __forceinline__ ClipView clip_view(const LogMelArgs &a, int b)
{
    ClipView v;
    if (a.srcOff) {
        v.L = max(ld_const(a.lengths, b), 0);
        v.x = a.src + ld_const(a.srcOff, b);
    } else {
        int start;
        v.L = row_region(a.begin, a.lengths, b, a.srcStride, start);   // never read past the clip's row
        v.x = a.src + static_cast<size_t>(b) * a.srcStride + start;
    }
    return clip_view_out(a, v, b);
}

// Address of mel row m, frame f of a clip's output.
__device__ __forceinline__ float *out_at(const LogMelArgs &a, const ClipView &v, int m, int f)
{
    return a.ntf ? v.out + static_cast<size_t>(f) * v.hs + m : v.out + static_cast<size_t>(m) * v.hs + f;
}

// Folds a thread's running maximum of stored mel power into the clip's (whole wave active).
__device__ __forceinline__ void clip_max_update(const LogMelArgs &a, int b, float mx)
{
    mx = wave_max(mx);
    if ((threadIdx.x & 63) == 0) atomicMax(&a.clipMax[b], __float_as_int(mx));
}

// Pre-emphasised sample at signal index j of a clip of length L, with the spectrogram's padding
// rule applied to out-of-range indices (the chain pads the pre-emphasised signal).
__device__ __forceinline__ float pe_sample(const float *x, int L, int j, float c, int border, int reflect)
{
    if (L <= 0) return 0.0f;
    if (j < 0 || j >= L) {
        if (!reflect) return 0.0f;
        j = idx_reflect(j, L);
    }
    float prev;
    if (j == 0) prev = (border == CLAMP) ? x[0] : (border == REFLECT && L > 1 ? x[1] : 0.0f);
    else prev = x[j - 1];
    float p = c * prev;
    return x[j] - p;
}

// Finds the clip that owns `tile` (largest b with tileStart[b] <= tile).
__device__ __forceinline__ int owner_clip(const int *tileStart, int B, int tile)
{
    int lo = 0, hi = B;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (tileStart[mid] <= tile) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ float db_fast(float v, const LogMelArgs &a)
{
    float r = v * a.invRef;
    float m = (a.minRatio < r) ? r : a.minRatio;
    return a.mult * log2f(m);
}

__device__ __forceinline__ float db_exact(float v, const LogMelArgs &a)
{
    float r = v * a.invRef;
    float m = (a.minRatio < r) ? r : a.minRatio;
    return a.mult * static_cast<float>(log2(static_cast<double>(m)));
}

// ------------------------------------------------------------------------- planner
// Single block: tiles per clip = ceil(F_b / FT) with F_b = min(frames(L_b), hStride); exclusive
// prefix sum into tileStart[0..B], total in tileStart[B]; with tileRec (strided layout only), one
// record per tile (clip, first frame, region length, region start).
__global__ void __launch_bounds__(1024)
plan_tiles_kernel(const int *__restrict__ lengths, const int *__restrict__ begin, int B, uint32_t srcStride,
                  uint32_t hStride, int W, int H, int center, int FT, int packed, int *__restrict__ tileStart,
                  int *__restrict__ clipMax, int4 *__restrict__ tileRec)
{
    if (clipMax)  // the fused kernel raises these with atomicMax (float bits; stored power is never below -FLT_MAX)
        for (int b = threadIdx.x; b < B; b += blockDim.x) clipMax[b] = __float_as_int(-3.40282347e38f);
    __shared__ int waveSum[16];
    __shared__ int carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < B; base += 1024) {
        const int b = base + tid;
        int n = 0, L = 0, start = 0;
        if (b < B) {
            L = packed ? lengths[b] : row_region(begin, lengths, b, srcStride, start);
            int F = max(0, num_windows(max(L, 0), W, H, center != 0));
            if (!packed) F = min(F, static_cast<int>(hStride));
            n = (F + FT - 1) / FT;
        }
        int v = n;  // inclusive wave scan
        for (int o = 1; o < 64; o <<= 1) {
            int u = __shfl_up(v, o);
            if (lane >= o) v += u;
        }
        if (lane == 63) waveSum[wave] = v;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; w++) before += waveSum[w];
        if (b < B) {
            const int ts = carry + before + v - n;
            tileStart[b] = ts;
            if (tileRec)
                for (int k = 0; k < n; k++) tileRec[ts + k] = make_int4(b, k * FT, L, start);
        }
        __syncthreads();
        if (tid == 1023) carry += before + v;
        __syncthreads();
    }
    if (tid == 0) tileStart[B] = carry;
}

// ------------------------------------------------------------------------- shared steps
// Stage the segment of frames [f0, f0 + FT) into LDS (`span` samples); only the frames below F
// are read from the clip, the rest of the span is zero.  Returns the block-wide max |sample| when
// `wantMax` (for the fp16 scaling), else 0.
template <bool wantMax>
__device__ float stage_segment(float *seg, const LogMelArgs &a, const float *x, int L, int f0, int F, int span,
                               float *red)
{
    const int off = a.center ? a.W / 2 : 0;
    const int s0 = f0 * a.H - off;
    const int nfr = max(0, F - f0);
    const int valid = nfr > 0 ? min(span, (nfr - 1) * a.H + a.W) : 0;
    float mx = 0.0f;
    for (int i = threadIdx.x; i < span; i += kThreads) {
        float v = i < valid ? pe_sample(x, L, s0 + i, a.coeff, a.border, a.reflect) : 0.0f;
        seg[seg_pos(i, a.H)] = v;
        if (wantMax) mx = fmaxf(mx, fabsf(v));
    }
    if (!wantMax) return 0.0f;
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    float r = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    return r;
}

// Mel projection + dB by the exact FP32 gather (same order as mel_filter_bank.hip), from the
// power tile pw[FT][pws].
__device__ void mel_gather_store(const float *pw, int pws, const LogMelArgs &a, const ClipView &v, int f0, int FT,
                                 bool exactLog)
{
    const int nOut = a.M * FT;
    float mx = -3.40282347e38f;
    for (int e = threadIdx.x; e < nOut; e += kThreads) {
        const int m = e / FT, f = e % FT;
        if (f0 + f >= v.F) continue;
        const float norm = a.normFactors[m];
        const int k0 = a.filterBins[2 * m], k1 = a.filterBins[2 * m + 1];
        float acc = 0.0f;
        for (int k = k0; k < k1; k++) {
            const float wd = a.weightsDown[k];
            float w = (a.intervals[k] == m) ? 1.0f - wd : wd;
            if (a.normalize) w = w * norm;
            const float p = pw[f * pws + k] * w;
            acc = acc + p;
        }
        if (a.clipMax) {
            *out_at(a, v, m, f0 + f) = acc;
            mx = fmaxf(mx, acc);
        } else {
            *out_at(a, v, m, f0 + f) = exactLog ? db_exact(acc, a) : db_fast(acc, a);
        }
    }
    if (a.clipMax) clip_max_update(a, v.b, mx);
}

// Second pass when the dB reference is the clip maximum (cpu/kernel/to_decibels.cpp:58-75): power
// stored by the fused kernel becomes mult * log2(max(minRatio, x / max)), 1 / max read as 1 for a
// zero maximum.  blockIdx.y strides over the clips.
__global__ void __launch_bounds__(256) db_from_clip_max_kernel(LogMelArgs a, int exactLog)
{
    for (int b = blockIdx.y; b < a.B; b += gridDim.y) {
        const ClipView v = clip_view(a, b);
        const float ref = __int_as_float(a.clipMax[b]);
        LogMelArgs c = a;
        c.invRef = ref != 0.0f ? 1.0f / ref : 1.0f;
        const long long n = static_cast<long long>(a.M) * v.F;
        for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
             e += static_cast<long long>(gridDim.x) * blockDim.x) {
            const int m = static_cast<int>(e / v.F), f = static_cast<int>(e % v.F);
            float *o = out_at(a, v, m, f);
            *o = exactLog ? db_exact(*o, c) : db_fast(*o, c);
        }
    }
}

// ------------------------------------------------------------------------- FP32 kernel
// FT = 32 frames per tile; each thread owns 4 frames x 8 column pairs (cg, cg + 32, ...) of a
// 512-column pass; the K loop order (t ascending, fmaf) and every later step match spectrogram.hip /
// mel_filter_bank.hip / to_decibels.hip exactly, so the output equals the 4-call chain bit for bit.
constexpr int kFT32 = 32;
constexpr int kKC = 16;
constexpr int kCC = 512;

__global__ void __launch_bounds__(kThreads, 1)
logmel_fp32_kernel(LogMelArgs a, int segFloats, int bFloats, int pwFloats, int aliasPw)
{
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ __attribute__((aligned(16))) float sA[kKC][kFT32];
    float *seg = lds;
    float *sB = lds + segFloats;
    float *pw = aliasPw ? sB : sB + bFloats;
    const int pws = a.KP + kSkew;
    const int tid = threadIdx.x;
    const int cg = tid & 31, fg = tid >> 5;
    const int T = a.tileStart[a.B];
    const int nPasses = (a.cols + kCC - 1) / kCC;
    const int span = (kFT32 - 1) * a.H + a.W;
    for (int tile = blockIdx.x; tile < T; tile += gridDim.x) {
        const int b = owner_clip(a.tileStart, a.B, tile);
        const int f0 = (tile - a.tileStart[b]) * kFT32;
        const ClipView v = clip_view(a, b);
        stage_segment<false>(seg, a, v.x, v.L, f0, v.F, span, nullptr);
        __syncthreads();
        for (int pass = 0; pass < nPasses; pass++) {
            const int cBase = pass * kCC;
            float acc[4][16];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 16; j++) acc[i][j] = 0.0f;
            for (int t0 = 0; t0 < a.W; t0 += kKC) {
                for (int e = tid; e < kKC * kCC; e += kThreads) {
                    int tt = e / kCC, cc = e % kCC;
                    int t = t0 + tt, c = cBase + cc;
                    sB[e] = (t < a.W && c < a.cols) ? a.basis[static_cast<size_t>(t) * a.cols + c] : 0.0f;
                }
                for (int e = tid; e < kKC * kFT32; e += kThreads) {  // the frames' samples of this chunk
                    const int tt = e / kFT32, ff = e % kFT32, t = t0 + tt;
                    sA[tt][ff] = (t < a.W) ? seg[seg_pos(ff * a.H + t, a.H)] : 0.0f;
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
            // power of this pass into pw[f][k]
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int f = fg * 4 + i;
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const float re = acc[i][2 * j], im = acc[i][2 * j + 1];
                    const int p = pass * (kCC / 2) + cg + 32 * j;
                    if (p == 0) {
                        float v = re * re, nq = im * im;
                        if (a.power == 1) { v = sqrtf(v); nq = sqrtf(nq); }
                        pw[f * pws + 0] = v;
                        if (!(a.N & 1)) pw[f * pws + a.N / 2] = nq;
                    } else if (p < a.K && !(!(a.N & 1) && p == a.N / 2)) {
                        float v = re * re + im * im;
                        if (a.power == 1) v = sqrtf(v);
                        pw[f * pws + p] = v;
                    }
                }
            }
            __syncthreads();
        }
        mel_gather_store(pw, pws, a, v, f0, kFT32, true);
        __syncthreads();
    }
}

// ------------------------------------------------------------------------- MFMA kernel
enum { kPrecBF16x3 = 1, kPrecFP16x3 = 2, kPrecBF16 = 3, kPrecBF16x6 = 4 };
// terms of each split DFT operand: hi, lo (and mid between them for the three-term bf16 split)
constexpr int split_terms(int prec) { return prec == kPrecBF16x6 ? 3 : 2; }

__device__ __forceinline__ void split_bf16(const float *v, bf16x8 &hi, bf16x8 &lo)
{
#pragma unroll
    for (int j = 0; j < 8; j++) {
        __bf16 h = static_cast<__bf16>(v[j]);
        hi[j] = h;
        lo[j] = static_cast<__bf16>(v[j] - static_cast<float>(h));
    }
}

__device__ __forceinline__ void split_f16(const float *v, float s, f16x8 &hi, f16x8 &lo)
{
#pragma unroll
    for (int j = 0; j < 8; j++) {
        float u = v[j] * s;
        _Float16 h = static_cast<_Float16>(u);
        hi[j] = h;
        lo[j] = static_cast<_Float16>(u - static_cast<float>(h));
    }
}

template <int PREC>
__device__ __forceinline__ void split8(const float *v, float s, u16x8 &hi, u16x8 &lo)
{
    if (PREC == kPrecFP16x3) {
        f16x8 h, l;
        split_f16(v, s, h, l);
        hi = __builtin_bit_cast(u16x8, h);
        lo = __builtin_bit_cast(u16x8, l);
    } else {
        bf16x8 h, l;
        split_bf16(v, h, l);
        hi = __builtin_bit_cast(u16x8, h);
        lo = __builtin_bit_cast(u16x8, l);
    }
}

// acc += A * B over the split operands: hi*hi + hi*lo + lo*hi (one product for kPrecBF16).
template <int PREC>
__device__ __forceinline__ void mfma3(f32x16 &acc, const u16x8 &ah, const u16x8 &al, const u16x8 &bh, const u16x8 &bl)
{
    if (PREC == kPrecFP16x3) {
        const f16x8 a0 = __builtin_bit_cast(f16x8, ah), a1 = __builtin_bit_cast(f16x8, al);
        const f16x8 b0 = __builtin_bit_cast(f16x8, bh), b1 = __builtin_bit_cast(f16x8, bl);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, b1, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, b0, acc, 0, 0, 0);
    } else {
        const bf16x8 a0 = __builtin_bit_cast(bf16x8, ah), a1 = __builtin_bit_cast(bf16x8, al);
        const bf16x8 b0 = __builtin_bit_cast(bf16x8, bh), b1 = __builtin_bit_cast(bf16x8, bl);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
        if (PREC == kPrecBF16x3) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc, 0, 0, 0);
        }
    }
}

// acc += A * B over three-term bf16 operands (hi, mid, lo: each term about 2^-8 of the one before):
// the six products down to 2^-16 relative -- hi*hi, hi*mid, mid*hi, hi*lo, lo*hi, mid*mid.
__device__ __forceinline__ void mfma6(f32x16 &acc, const u16x8 (&a)[3], const u16x8 (&b)[3])
{
    const bf16x8 a0 = __builtin_bit_cast(bf16x8, a[0]), a1 = __builtin_bit_cast(bf16x8, a[1]),
                 a2 = __builtin_bit_cast(bf16x8, a[2]);
    const bf16x8 b0 = __builtin_bit_cast(bf16x8, b[0]), b1 = __builtin_bit_cast(bf16x8, b[1]),
                 b2 = __builtin_bit_cast(bf16x8, b[2]);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b2, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a2, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
}

// The products of one split operand pair for any precision (NT terms per operand).
template <int PREC, int NT>
__device__ __forceinline__ void mfma_split(f32x16 &acc, const u16x8 (&a)[NT], const u16x8 (&b)[NT])
{
    if constexpr (PREC == kPrecBF16x6) mfma6(acc, a, b);
    else mfma3<PREC>(acc, a[0], a[1], b[0], b[1]);
}

// Largest b with tileStart[b] <= tile, by a 64-ary search over the wave (two rounds of loads for a
// few hundred clips).  tileStart is non-decreasing, so the lanes whose candidate passes form a
// prefix.  Call with the whole wave active.
__device__ __forceinline__ int owner_clip_wave(const int *tileStart, int B, int tile, int lane)
{
    int lo = 0, hi = B;
    while (hi - lo > 1) {
        const int step = (hi - lo + 63) / 64;
        const int cand = lo + lane * step;
        const bool ok = cand < hi && tileStart[cand] <= tile;
        const int cnt = __popcll(__ballot(ok));
        lo = lo + (cnt - 1) * step;
        hi = min(hi, lo + step);
    }
    return lo;
}

// One tile of the work list: its clip, first frame, and where its segment [s0, s0 + span) sits.
struct TileInfo {
    ClipView v;
    int f0, s0, valid;   // valid: samples of the span that belong to a frame below F (the rest are zero)
    int iLo, iHi;        // first and last interior group start (iHi < iLo: none)
    bool aligned;        // clip samples at s0 start 16-byte aligned
};

__device__ __forceinline__ TileInfo tile_info(const LogMelArgs &a, int tile, int FT, int span, int lane)
{
    TileInfo t;
    if (a.blockRec) {
        // the block's own record k, written by this block earlier in this launch: a plain global load
        // (the stores and a workgroup barrier precede it, so it sees them), made wave-uniform.  Not
        // through the constant address space, whose loads the compiler may treat as invariant for
        // the whole launch and move above the stores.
        const int k = (tile - static_cast<int>(blockIdx.x)) / static_cast<int>(gridDim.x);
        const int4 g = a.blockRec[static_cast<size_t>(blockIdx.x) * a.kPad + k];
        int4 r;
        r.x = __builtin_amdgcn_readfirstlane(g.x);
        r.y = __builtin_amdgcn_readfirstlane(g.y);
        r.z = __builtin_amdgcn_readfirstlane(g.z);
        r.w = __builtin_amdgcn_readfirstlane(g.w);
        ClipView v;
        v.L = r.z;
        v.x = a.src + static_cast<size_t>(r.x) * a.srcStride + r.w;
        t.v = clip_view_out(a, v, r.x);
        t.f0 = r.y;
    } else if (a.tileRec) {
        // one uniform 16-byte load instead of the search and the clip's own loads
        typedef int i32x4 __attribute__((ext_vector_type(4)));
        const i32x4 r = ld_const(reinterpret_cast<const i32x4 *>(a.tileRec), tile);
        ClipView v;
        v.L = r.z;
        v.x = a.src + static_cast<size_t>(r.x) * a.srcStride + r.w;
        t.v = clip_view_out(a, v, r.x);
        t.f0 = r.y;
    } else {
        const int b = owner_clip_wave(a.tileStart, a.B, tile, lane);
        t.v = clip_view(a, b);
        t.f0 = (tile - ld_const(a.tileStart, b)) * FT;
    }
    t.s0 = t.f0 * a.H - (a.center ? a.W / 2 : 0);
    const int nfr = max(0, t.v.F - t.f0);
    t.valid = (nfr > 0 && t.v.L > 0) ? min(span, (nfr - 1) * a.H + a.W) : 0;
    t.aligned = ((reinterpret_cast<uintptr_t>(t.v.x) + 4 * static_cast<intptr_t>(t.s0)) & 15) == 0;
    // interior group starts i (multiples of 4): s0 + i - 1 >= 0, s0 + i + 3 < L, i + 3 < valid
    t.iLo = max(0, (1 - t.s0 + 3) & ~3);
    const int hi = min(t.v.L - 4 - t.s0, t.valid - 4);
    t.iHi = hi < 0 ? -4 : (hi & ~3);
    return t;
}

// Staging works on groups of four samples (group g, samples 4g .. 4g+3 of the span, to thread
// g % blockDim).  An interior group -- its samples and the one before them inside the clip, all in
// the valid part -- is pre-emphasised from plain loads; the others (clip edges: reflect/zero
// padding, the border rule, samples past the last frame) sample by sample.  The interior groups
// of a thread's first kPF rounds are loaded ahead, into registers, while the previous tile's mel
// product runs.
constexpr int kPF = 11;
constexpr int kBlockMfma = 256;   // threads of a fused MFMA block (the launch always uses this)
struct SegPrefetch {
    float4 v[kPF];
    float prev[kPF];
};

__device__ __forceinline__ bool interior_group(const TileInfo &t, int i) { return i >= t.iLo && i <= t.iHi; }

__device__ __forceinline__ void load_group(const TileInfo &t, int i, float4 &x4, float &prev)
{
    const float *x = t.v.x + t.s0 + i;
    if (t.aligned) x4 = *reinterpret_cast<const float4 *>(x);
    else x4 = make_float4(x[0], x[1], x[2], x[3]);
    prev = x[-1];
}

// Fresh contents: the registers of the previous prefetch must not stay live across a tile whose
// successor loads nothing.
__device__ __forceinline__ void seg_prefetch_none(SegPrefetch &p)
{
#pragma unroll
    for (int q = 0; q < kPF; q++) {
        // an empty asm statement defines the registers without an instruction (zero-filling them
        // would cost 55 moves per tile; the values are never read)
        float4 u;
        float w;
        __asm__ volatile("" : "=v"(u.x), "=v"(u.y), "=v"(u.z), "=v"(u.w), "=v"(w));
        p.v[q] = u;
        p.prev[q] = w;
    }
}

// Straight-line loads (a group outside the interior reads the nearest interior group instead, and
// its values are never used), so the compiler keeps them in registers while the mel product runs.
__device__ __forceinline__ void seg_prefetch(SegPrefetch &p, const TileInfo &t, int span, int tid)
{
    if (t.iHi < t.iLo) {
        seg_prefetch_none(p);
        return;
    }
    const int nt = kBlockMfma;
#pragma unroll
    for (int q = 0; q < kPF; q++) {
        const int i = min(max(4 * (tid + q * nt), t.iLo), t.iHi);
        load_group(t, i, p.v[q], p.prev[q]);
    }
}

__device__ __forceinline__ float4 pe_group(const float4 x, float prev, float c)
{
    float4 y;
    float t;
    t = c * prev; y.x = x.x - t;
    t = c * x.x;  y.y = x.y - t;
    t = c * x.y;  y.z = x.z - t;
    t = c * x.z;  y.w = x.w - t;
    return y;
}

__device__ __forceinline__ float pe_sample_sel(const float *x, int L, int j, float c, int border, int reflect);

// Edge group: sample by sample, zero past the valid part.
__device__ __forceinline__ float4 edge_group(const LogMelArgs &a, const TileInfo &t, int i)
{
    float y[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
        y[k] = (i + k < t.valid) ? pe_sample_sel(t.v.x, t.v.L, t.s0 + i + k, a.coeff, a.border, a.reflect) : 0.0f;
    return make_float4(y[0], y[1], y[2], y[3]);
}

// Writes the tile's pre-emphasised segment into LDS (skewed); returns the thread's max |y|.
__device__ __forceinline__ float seg_store(float *seg, const SegPrefetch &p, const LogMelArgs &a, const TileInfo &t,
                                           int span, float invH, int tid)
{
    const int groups = (span + 3) / 4, nt = kBlockMfma;
    float mx = 0.0f;
    auto put = [&](int i, float4 y) {
        *reinterpret_cast<float4 *>(&seg[seg_pos_r(i, invH)]) = y;
        mx = fmaxf(mx, fmaxf(fmaxf(fabsf(y.x), fabsf(y.y)), fmaxf(fabsf(y.z), fabsf(y.w))));
    };
#pragma unroll
    for (int q = 0; q < kPF; q++) {  // interior groups of the prefetched rounds
        const int g = tid + q * nt, i = 4 * g;
        if (g < groups && interior_group(t, i)) put(i, pe_group(p.v[q], p.prev[q], a.coeff));
    }
    // edge groups, and rounds past the prefetch (none when every group is interior and prefetched)
    const bool allPrefetched = t.iLo == 0 && t.iHi >= 4 * (groups - 1) && groups <= kPF * nt;
    for (int g = allPrefetched ? groups : tid, q = 0; g < groups; g += nt, q++) {
        const int i = 4 * g;
        const bool in = interior_group(t, i);
        if (in && q < kPF) continue;
        float4 y;
        if (in) {
            float4 x4;
            float prev;
            load_group(t, i, x4, prev);
            y = pe_group(x4, prev, a.coeff);
        } else {
            y = edge_group(a, t, i);
        }
        put(i, y);
    }
    return mx;
}

// pe_sample with the index and border rule resolved before the loads (the reflect map only for
// samples outside the clip), so a thread can have several samples' loads in flight.  Same
// arithmetic as pe_sample.
__device__ __forceinline__ float pe_sample_sel(const float *x, int L, int j, float c, int border, int reflect)
{
    int jj = j;
    bool zero = false;
    if (j < 0 || j >= L) {
        if (reflect) jj = idx_reflect(j, L);
        else zero = true;
        jj = min(max(jj, 0), L - 1);
    }
    int pi = jj - 1;
    bool pz = false;
    if (pi < 0) {
        pi = (border == CLAMP || L < 2) ? 0 : 1;
        pz = !(border == CLAMP || (border == REFLECT && L > 1));
    }
    const float xv = x[jj];
    const float pv = pz ? 0.0f : x[pi];
    const float p = c * pv;
    const float y = xv - p;
    return zero ? 0.0f : y;
}

// Two-element hi+lo split, packed as the adjacent halves of an MFMA operand (element j in the low
// 16 bits), with the same rounding as split8.
template <int PREC>
__device__ __forceinline__ void split2(float v0, float v1, float s, uint32_t &hi, uint32_t &lo)
{
    if (PREC == kPrecFP16x3) {
        // one mixed-precision FMA per half: hi = f16(v*s), lo = f16(v*s - hi).  s is a power of
        // two, so v*s is exact in FP32 and so is v*s - hi (it has at most 13 significant bits):
        // each result is rounded once, to nearest even, as the conversions round it
        uint32_t h, l;
        __asm__("v_fma_mixlo_f16 %0, %1, %2, 0" : "=v"(h) : "v"(v0), "v"(s));
        __asm__("v_fma_mixhi_f16 %0, %1, %2, 0" : "+v"(h) : "v"(v1), "v"(s));
        __asm__("v_fma_mixlo_f16 %0, %1, %2, -%3 op_sel_hi:[0,0,1]" : "=v"(l) : "v"(v0), "v"(s), "v"(h));
        __asm__("v_fma_mixhi_f16 %0, %1, %2, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(l) : "v"(v1), "v"(s), "v"(h));
        hi = h;
        lo = l;
    } else {
        const __bf16 h0 = static_cast<__bf16>(v0), h1 = static_cast<__bf16>(v1);
        const __bf16 l0 = static_cast<__bf16>(v0 - static_cast<float>(h0));
        const __bf16 l1 = static_cast<__bf16>(v1 - static_cast<float>(h1));
        hi = __builtin_bit_cast(uint16_t, h0) | (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, h1)) << 16);
        lo = __builtin_bit_cast(uint16_t, l0) | (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, l1)) << 16);
    }
}

// Two-element three-term bf16 split: hi = bf16(v), mid = bf16(v - hi), lo = bf16(v - hi - mid)
// (each difference exact in FP32), packed as split2 packs them.
__device__ __forceinline__ void split3_bf16(float v0, float v1, uint32_t &hi, uint32_t &mid, uint32_t &lo)
{
    const __bf16 h0 = static_cast<__bf16>(v0), h1 = static_cast<__bf16>(v1);
    const float r0 = v0 - static_cast<float>(h0), r1 = v1 - static_cast<float>(h1);
    const __bf16 m0 = static_cast<__bf16>(r0), m1 = static_cast<__bf16>(r1);
    const __bf16 l0 = static_cast<__bf16>(r0 - static_cast<float>(m0));
    const __bf16 l1 = static_cast<__bf16>(r1 - static_cast<float>(m1));
    auto pack = [](__bf16 x, __bf16 y) {
        return __builtin_bit_cast(uint16_t, x) | (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, y)) << 16);
    };
    hi = pack(h0, h1);
    mid = pack(m0, m1);
    lo = pack(l0, l1);
}

// The split of two elements into the NT terms of precision PREC, term t of the pair in w[t].
template <int PREC>
__device__ __forceinline__ void split_terms2(float v0, float v1, float s, uint32_t (&w)[split_terms(PREC)])
{
    if constexpr (PREC == kPrecBF16x6) split3_bf16(v0, v1, w[0], w[1], w[2]);
    else split2<PREC>(v0, v1, s, w[0], w[1]);
}

// Folded, split DFT operand of SPB 16-row contraction steps starting at step s, built into buffer
// `buf` of the slice ring while the waves run the previous steps' products: slot
// ((buf * 2NT + arr) * 2SPB + g) * FT + f holds 8 halves (16 bytes) of frame f, rows 16s + 8g .. +7,
// for arr = t the e operand's term t, NT + t the o operand's (NT = 2: hi, lo; 3: hi, mid, lo) --
// the lane order of a 32x32x16 MFMA fragment.  A work
// unit forms four rows of one frame (one 8-byte half of a slot).  q4 gathers the thread's share of
// bin N/4: re += e_j w_j c_j, im += o_j w_j s_j with (c_j, s_j) = sqrt(2) (cos, sin)(pi (2j+1)/4),
// i.e. c = +1 -1 -1 +1 and s = +1 +1 -1 -1 over j mod 4 (the sqrt(2)/2 factors are applied to the
// power).
constexpr int kSPB = 2;   // contraction steps per barrier

__device__ __forceinline__ void q4_accumulate(float2 &q4, float e0, float e1, float e2, float e3, float o0, float o1,
                                              float o2, float o3, float4 w)
{
    q4.x = fmaf(e0, w.x, q4.x);
    q4.x = fmaf(e1, -w.y, q4.x);
    q4.x = fmaf(e2, -w.z, q4.x);
    q4.x = fmaf(e3, w.w, q4.x);
    q4.y = fmaf(o0, w.x, q4.y);
    q4.y = fmaf(o1, w.y, q4.y);
    q4.y = fmaf(o2, -w.z, q4.y);
    q4.y = fmaf(o3, -w.w, q4.y);
}

template <int PREC, int FT>
__device__ __forceinline__ void prep_step(u16x8 *ring, const float *seg, const float *wHalf, const LogMelArgs &a,
                                          int s, int buf, float segScale, float invH, float2 &q4, int tid)
{
    constexpr int GB = 2 * kSPB;                         // 8-row groups per buffer
    const int half = a.W / 2;
    uint2 *dst = reinterpret_cast<uint2 *>(ring);
    for (int u = tid; u < 2 * GB * FT; u += kBlockMfma) {
        const int f = u % FT, qd = u / FT;
        const int j = 16 * s + 4 * qd;                 // rows j .. j+3 (j a multiple of 4)
        float e0 = 0.0f, e1 = 0.0f, e2 = 0.0f, e3 = 0.0f, o0 = 0.0f, o1 = 0.0f, o2 = 0.0f, o3 = 0.0f;
        if (j < half) {
            const int fb = f * (a.H + kSkew);           // seg_pos(f*H + t) = f*(H + skew) + seg_pos(t)
            const float4 up = *reinterpret_cast<const float4 *>(&seg[fb + seg_pos_r(half + j, invH)]);
            const float4 dn = *reinterpret_cast<const float4 *>(&seg[fb + seg_pos_r(half - 4 - j, invH)]);  // rows j+3..j
            e0 = up.x + dn.w; o0 = up.x - dn.w;
            e1 = up.y + dn.z; o1 = up.y - dn.z;
            e2 = up.z + dn.y; o2 = up.z - dn.y;
            e3 = up.w + dn.x; o3 = up.w - dn.x;
            q4_accumulate(q4, e0, e1, e2, e3, o0, o1, o2, o3, *reinterpret_cast<const float4 *>(&wHalf[j]));
        }
        constexpr int NT = split_terms(PREC);
        uint32_t e01[NT], e23[NT], o01[NT], o23[NT];
        split_terms2<PREC>(e0, e1, segScale, e01);
        split_terms2<PREC>(e2, e3, segScale, e23);
        split_terms2<PREC>(o0, o1, segScale, o01);
        split_terms2<PREC>(o2, o3, segScale, o23);
        const int slot = (buf * 2 * NT * GB + (qd >> 1)) * FT + f, hh = qd & 1;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            dst[(slot + t * GB * FT) * 2 + hh] = make_uint2(e01[t], e23[t]);
            dst[(slot + (NT + t) * GB * FT) * 2 + hh] = make_uint2(o01[t], o23[t]);
        }
    }
}

// The same work as prep_step for one unit u of the 2*kSPB*2*FT units of a step pair, split so
// the loads can be issued ahead of the products they overlap: prep_load reads the segment and
// window, prep_finish forms, splits and stores.  Rows past W/2 take the samples of the last real
// rows (finite values that meet all-zero basis rows) and the four zeros stored after the window
// half (so they add nothing to the bin N/4 sums): no masking.
struct PrepIn {
    float4 up, dn, w;
};

template <int FT>
__device__ __forceinline__ void prep_load(PrepIn &p, const float *seg, const float *wHalf, const LogMelArgs &a, int s,
                                          int u, float invH)
{
    const int f = u % FT, qd = u / FT;
    const int half = a.W / 2;
    const int j = 16 * s + 4 * qd;
    const int jc = min(j, half - 4);
    const int fb = f * (a.H + kSkew);
    p.up = *reinterpret_cast<const float4 *>(&seg[fb + seg_pos_r(half + jc, invH)]);
    p.dn = *reinterpret_cast<const float4 *>(&seg[fb + seg_pos_r(half - 4 - jc, invH)]);
    p.w = *reinterpret_cast<const float4 *>(&wHalf[min(j, half)]);
}

template <int PREC, int FT>
__device__ __forceinline__ void prep_finish(u16x8 *ring, const PrepIn &p, int u, int buf, float segScale, float2 &q4)
{
    constexpr int GB = 2 * kSPB;
    const int f = u % FT, qd = u / FT;
    const float4 up = p.up, dn = p.dn, w = p.w;
    const float e0 = up.x + dn.w, o0 = up.x - dn.w;
    const float e1 = up.y + dn.z, o1 = up.y - dn.z;
    const float e2 = up.z + dn.y, o2 = up.z - dn.y;
    const float e3 = up.w + dn.x, o3 = up.w - dn.x;
    q4_accumulate(q4, e0, e1, e2, e3, o0, o1, o2, o3, w);   // zero w past W/2 leaves q4 unchanged
    constexpr int NT = split_terms(PREC);
    uint32_t e01[NT], e23[NT], o01[NT], o23[NT];
    split_terms2<PREC>(e0, e1, segScale, e01);
    split_terms2<PREC>(e2, e3, segScale, e23);
    split_terms2<PREC>(o0, o1, segScale, o01);
    split_terms2<PREC>(o2, o3, segScale, o23);
    uint2 *dst = reinterpret_cast<uint2 *>(ring);
    const int slot = (buf * 2 * NT * GB + (qd >> 1)) * FT + f, hh = qd & 1;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        dst[(slot + t * GB * FT) * 2 + hh] = make_uint2(e01[t], e23[t]);
        dst[(slot + (NT + t) * GB * FT) * 2 + hh] = make_uint2(o01[t], o23[t]);
    }
}

// Scheduling hint for the region before it: NMFMA products, each followed by up to three vector
// ALU instructions (the operand build of the next step pair runs while the products execute).
template <int NMFMA>
__device__ __forceinline__ void interleave_valu()
{
#pragma unroll
    for (int i = 0; i < NMFMA; i++) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // MFMA
        __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);   // VALU
    }
}

// bf16 hi+lo split of one power value into the two planes the mel product reads (same rounding
// as split_bf16).
__device__ __forceinline__ void put_power(uint16_t *pwH, uint16_t *pwL, int idx, float p)
{
    const __bf16 h = static_cast<__bf16>(p);
    const __bf16 l = static_cast<__bf16>(p - static_cast<float>(h));
    pwH[idx] = __builtin_bit_cast(uint16_t, h);
    pwL[idx] = __builtin_bit_cast(uint16_t, l);
}

// A basis fragment with its odd rows negated (rows j..j+7 of one lane: the odd j are the high
// halves of its four dwords), the operand of the mirror bins N/2 - k.
__device__ __forceinline__ u16x8 flip_odd(const u16x8 v)
{
    // one 32-bit XOR per dword (written out: left to itself the compiler splits it into 16-bit
    // pieces, two instructions per dword)
    const u32x4 w = __builtin_bit_cast(u32x4, v);
    u32x4 r;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        uint32_t x;
        __asm__("v_xor_b32 %0, 0x80000000, %1" : "=v"(x) : "v"(w[i]));
        r[i] = x;
    }
    return __builtin_bit_cast(u16x8, r);
}

// bf16 hi and lo of four power values as two 8-byte plane stores.
__device__ __forceinline__ void put_power4(uint16_t *pwH, uint16_t *pwL, int idx, const float (&p)[4])
{
    // one packed conversion per pair for hi; the pair's FP32 values of hi are its low half shifted
    // up and its high half masked (no second conversion), then one packed conversion for lo
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    uint32_t h[2], l[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const uint32_t hp = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{p[2 * k], p[2 * k + 1]}, bf16x2));
        const float h0 = __uint_as_float(hp << 16), h1 = __uint_as_float(hp & 0xffff0000u);
        h[k] = hp;
        l[k] = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{p[2 * k] - h0, p[2 * k + 1] - h1}, bf16x2));
    }
    *reinterpret_cast<uint2 *>(&pwH[idx]) = make_uint2(h[0], h[1]);
    *reinterpret_cast<uint2 *>(&pwL[idx]) = make_uint2(l[0], l[1]);
}

// Power (|X| when SQRT) of the wave's units into the planes, in the scaled domain of the operands
// (the mel output undoes the scaling).  The DFT accumulators hold X^T: row =
// bin, column = frame, so a lane owns frame rt*32 + lane%32 and, per group of four registers,
// direct bins k0 .. k0+3 (k0 = ct*32 + 8q + 4(lane/32)): acc 0/1 give bin k (cos x e, sin x o),
// acc 2/3 bin N/2 - k (odd-negated sin x e, cos x o).  Bin k sits at plane position k, bin
// N/2 - k at N/2 + 3 - k, so both groups are 8-byte aligned runs (the mirror run reversed).
template <bool SQRT, int RTW, int CTM>
__device__ __forceinline__ void store_power(uint16_t *pwH, uint16_t *pwL, int pws2, const f32x16 (&acc)[RTW][CTM][4],
                                            int ct0, int rt0, int nct, int halfN, int lane)
{
#pragma unroll
    for (int r = 0; r < RTW; r++) {
#pragma unroll
        for (int ct = 0; ct < CTM; ct++) {
            if (ct0 + ct >= nct) continue;   // a repeated tile (uneven split over the waves)
            const int row = ((rt0 + r) * 32 + (lane & 31)) * pws2;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int k0 = (ct0 + ct) * 32 + 8 * q + 4 * (lane >> 5);
                float pd[4], pm[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float c = acc[r][ct][0][4 * q + i], sn = acc[r][ct][1][4 * q + i];
                    const float mc = acc[r][ct][2][4 * q + i], ms = acc[r][ct][3][4 * q + i];
                    float d = fmaf(sn, sn, c * c), m = fmaf(ms, ms, mc * mc);
                    if (SQRT) { d = sqrtf(d); m = sqrtf(m); }
                    pd[i] = d;
                    pm[3 - i] = m;
                }
                put_power4(pwH, pwL, row + k0, pd);
                put_power4(pwH, pwL, row + halfN - k0, pm);
            }
        }
    }
}

// Mel work list of one wave: entries (16-filter x 16-frame tile, 32-bin step of its band), packed
// as tile << 8 | step; every tile has at least one step.  Capacity per wave:
constexpr int kMelSeq = 320;
// Mel weight fragments in flight per wave (work-list entries loaded ahead).
constexpr int kWA = 2;
static_assert(kWA >= 2, "the tile-end test reads entry e + 1 from the next slot");

// The planner's work inside a block: tiles per clip (plan_tiles_kernel's count) and their
// exclusive prefix sum into ts[0..B] (ts[B] = total), ws[0..3] scratch; returns the total.  Strided
// layout only.
__device__ int plan_tiles_block(const LogMelArgs &a, int FT, int *ts, int *ws)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int carry = 0;
    for (int base = 0; base < a.B; base += kBlockMfma) {
        const int b = base + tid;
        int n = 0;
        if (b < a.B) {
            int start;
            const int L = row_region(a.begin, a.lengths, b, a.srcStride, start);
            const int F = min(max(0, num_windows(max(L, 0), a.W, a.H, a.center != 0)), static_cast<int>(a.frameCap));
            n = (F + FT - 1) / FT;
        }
        int v = n;   // inclusive wave scan
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o);
            if (lane >= o) v += u;
        }
        if (lane == 63) ws[wave] = v;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wave; w++) before += ws[w];
        if (b < a.B) ts[b] = carry + before + v - n;
        carry += ws[0] + ws[1] + ws[2] + ws[3];
        __syncthreads();   // ws is reused by the next chunk
    }
    if (tid == 0) ts[a.B] = carry;
    __syncthreads();
    return carry;
}

// Fused MFMA kernel.  Block = 4 waves.  The DFT of a tile (FT frames = NRT row tiles of 32) runs
// over N/128 column tiles of 32 direct bins k < N/4, each giving bins k and N/2 - k; the
// (column tile, row tile) units are split evenly over the waves, CTM x RTW per wave, four 32x32
// accumulator tiles per unit.  Two blocks share a CU, so one block's staging and power/mel phases
// overlap the other's products.  Per tile: stage the pre-emphasised segment; per pair of 16-row
// steps of the folded contraction, the products run from one buffer of the operand ring while
// the block builds the next pair into the other (one barrier per pair); power in lane into bf16
// hi/lo planes; banded mel on MFMA, each weight fragment applied to every frame group of the
// tile; dB out.
template <int PREC, int FT, int CTM, int RTW>
__global__ void __launch_bounds__(kBlockMfma, PREC == kPrecBF16x6 ? 1 : 2) logmel_mfma_kernel(LogMelArgs a, int segFloats,
                                                                                             int wOff)
{
    constexpr int NT = split_terms(PREC);   // terms per split operand
    constexpr int NRT = FT / 32;
    constexpr int U = CTM * RTW;                         // units per wave
    constexpr int NW = 4;
    static_assert(U <= 2, "accumulator budget: eight 32x32 tiles per wave");
    static_assert(RTW <= NRT, "a wave's row tiles lie in one tile");
    extern __shared__ __attribute__((aligned(16))) float lds[];
    __shared__ float red[NW];
    __shared__ float2 q4Part[NW * 64];                  // bin N/4 partial sums (re, im)
    __shared__ uint16_t melSeq[NW][kMelSeq];
    __shared__ int melSeqLen[NW];
    __shared__ uint16_t melBand[16];                     // filter tile's band: first step << 8 | end step
    float *seg = lds;                                    // the power planes alias segment + ring after the DFT
    u16x8 *ring = reinterpret_cast<u16x8 *>(lds + segFloats);   // [2][2 NT][2 kSPB][FT] operand slices
    const int pws2 = a.KP + 8;                           // power plane row stride (halves; 16-byte skew)
    const int span = (FT - 1) * a.H + a.W;
    const int halfN = a.N / 2;
    const int nSteps = a.Wp / 16;
    const float invH = 1.0f / static_cast<float>(a.H);
    const int nMT = a.Mpad / 16;
    float *wH = lds + wOff;                              // window half (bin N/4 sums), past the aliased region
    for (int j = threadIdx.x; j < a.W / 2 + 4; j += kBlockMfma) wH[j] = j < a.W / 2 ? a.wHalf[j] : 0.0f;

    // mel work lists (the same for every tile): an entry (filter tile, half of the frame groups,
    // 32-bin step of the band) applies one weight fragment to FT/32 frame groups
    if (threadIdx.x < nMT) {
        const int mt = threadIdx.x;
        int kb = a.KP, ke = 0;
        for (int i = 0; i < 16; i++) {
            const int m = mt * 16 + i;
            if (m < a.M) {
                kb = min(kb, a.filterBins[2 * m]);
                ke = max(ke, a.filterBins[2 * m + 1]);
            }
        }
        // plane positions: bin b at b, or b + 3 above N/4
        const int Q = a.N / 4;
        if (kb < ke) {
            kb = kb <= Q ? kb : kb + 3;
            ke = (ke - 1 <= Q ? ke - 1 : ke + 2) + 1;
        }
        kb = kb >= ke ? 0 : kb / 32 * 32;        // an empty band still takes one (all-zero) step
        ke = max(min(a.KP, (ke + 31) / 32 * 32), kb + 32);
        melBand[mt] = static_cast<uint16_t>((kb / 32) << 8 | (ke / 32));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // work items: (filter tile, half of the tile's frame groups), costing the band's steps plus
        // one for the stores at the item's end; largest first, each to the least-loaded wave
        int load[NW] = {0, 0, 0, 0}, cnt[NW] = {0, 0, 0, 0};
        uint32_t taken = 0;
        const int nItems = 2 * nMT;
        for (int it = 0; it < nItems; it++) {
            int best = -1, bestCost = -1;
            for (int item = 0; item < nItems; item++) {
                const int mt = item >> 1;
                const int cost = (melBand[mt] & 255) - (melBand[mt] >> 8) + 1;
                if (!(taken >> item & 1) && cost > bestCost) { best = item; bestCost = cost; }
            }
            taken |= 1u << best;
            int w = 0;
            for (int q = 1; q < NW; q++) if (load[q] < load[w]) w = q;
            const int mt = best >> 1;
            // at most 2 * 16 tiles x 17 steps = 544 entries in all (numFilter <= 256, nfft <= 1024),
            // so the least-loaded wave stays below 544 / 4 + 17 <= kMelSeq; the bound is checked anyway
            for (int k = melBand[mt] >> 8; k < (melBand[mt] & 255) && cnt[w] < kMelSeq; k++)
                melSeq[w][cnt[w]++] = static_cast<uint16_t>((best << 8) | k);
            load[w] += bestCost;
        }
        for (int q = 0; q < NW; q++) melSeqLen[q] = cnt[q];
    }

    // the tile plan: computed here, in the segment region (free until the first tile is staged),
    // with this block's own tile records written out; or the planner's
    int T;
    if (a.blockRec) {
        int *ts = reinterpret_cast<int *>(lds);
        T = plan_tiles_block(a, FT, ts, ts + a.B + 1);
        for (int k = threadIdx.x; k < a.kPad; k += kBlockMfma) {
            const int tile = static_cast<int>(blockIdx.x) + k * static_cast<int>(gridDim.x);
            if (tile < T) {
                int lo = 0, hi = a.B;   // ts[lo] <= tile < ts[hi]: the largest such lo owns the tile
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (ts[mid] <= tile) lo = mid;
                    else hi = mid;
                }
                int start;
                const int L = row_region(a.begin, a.lengths, lo, a.srcStride, start);
                a.blockRec[static_cast<size_t>(blockIdx.x) * a.kPad + k] = make_int4(lo, (tile - ts[lo]) * FT, L, start);
            }
        }
        // the records are read back by this block's waves with ordinary global loads; the stores
        // are waited for explicitly before the barrier (once per launch)
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
        __syncthreads();   // records stored; the segment region is free again
    } else {
        T = ld_const(a.tileStart, a.B);
    }

    // the first tile's segment is fetched here; every later one while its predecessor's mel runs
    SegPrefetch pf;
    TileInfo cur;
    int tile = blockIdx.x;
    if (tile < T) {
        const int tid = threadIdx.x;
        cur = tile_info(a, tile, FT, span, tid & 63);
        seg_prefetch(pf, cur, span, tid);
    }
    for (; tile < T; tile += gridDim.x) {
        // thread-index values are re-derived from a fresh opaque copy in every phase, so none of
        // them is computed once and kept live across the others
        int tid = opaque(threadIdx.x), wave = tid >> 6, lane = tid & 63;
        const ClipView v = cur.v;
        const int f0 = cur.f0;

        // ---- stage the pre-emphasised segment; block max for the fp16 scale
        float mx = seg_store(seg, pf, a, cur, span, invH, tid);
        mx = wave_max(mx);
        if (lane == 0) red[wave] = mx;
        __syncthreads();
        mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        float segScale = 1.0f;
        // 2*max*scale < 2^11 (the exponent capped so segScale * basis scale stays finite for
        // near-zero segments)
        if (PREC == kPrecFP16x3 && mx > 0.0f) segScale = ldexpf(1.0f, min(9 - ilogbf(mx), 114));

        // ---- folded DFT
        tid = opaque(threadIdx.x); wave = tid >> 6; lane = tid & 63;
        const int nct = a.N / 128;                       // column tiles of direct bins
        const int ct0 = wave * U / NRT, rt0 = wave * U % NRT;
        f32x16 acc[RTW][CTM][4];
#pragma unroll
        for (int r = 0; r < RTW; r++)
#pragma unroll
            for (int ct = 0; ct < CTM; ct++)
#pragma unroll
                for (int q = 0; q < 4; q++)
#pragma unroll
                    for (int i = 0; i < 16; i++) acc[r][ct][q][i] = 0.0f;
        // basis fragments [buffer][col tile][hi, lo]: one column tile per wave double-buffers whole
        // steps; two refill each tile's fragments right after its own products (the other tile's
        // products cover the latency), which keeps the register budget.  A tile index past the
        // last (four waves over six tiles) repeats the last tile; its power is never stored.
        constexpr int NB = CTM == 1 ? 2 : 1;
        u16x8 bc[NB][CTM][NT], bs[NB][CTM][NT];
        const int NC = a.N / 2;
        // 32-bit byte offsets from the table bases (the tables stay far below 4 GB), so the loads
        // take a scalar base and a vector offset
        const char *bHi = reinterpret_cast<const char *>(a.bHi), *bLo = reinterpret_cast<const char *>(a.bLo);
        const char *bLo2 = reinterpret_cast<const char *>(a.bLo2);
        auto loadB = [&](int buf, int ct, int s) {
            const uint32_t c = min(ct0 + ct, nct - 1) * 32 + (lane & 31);
            const uint32_t oc = (static_cast<uint32_t>(2 * s + (lane >> 5)) * NC + c) * 16u;
            const uint32_t os = oc + static_cast<uint32_t>(NC / 2) * 16u;
            bc[buf][ct][0] = *reinterpret_cast<const u16x8 *>(bHi + oc);
            bs[buf][ct][0] = *reinterpret_cast<const u16x8 *>(bHi + os);
            if (PREC != kPrecBF16) {
                bc[buf][ct][1] = *reinterpret_cast<const u16x8 *>(bLo + oc);
                bs[buf][ct][1] = *reinterpret_cast<const u16x8 *>(bLo + os);
            }
            if constexpr (NT == 3) {
                bc[buf][ct][NT - 1] = *reinterpret_cast<const u16x8 *>(bLo2 + oc);
                bs[buf][ct][NT - 1] = *reinterpret_cast<const u16x8 *>(bLo2 + os);
            }
        };
        float2 q4 = make_float2(0.0f, 0.0f);
        constexpr int GB = 2 * kSPB;
        prep_step<PREC, FT>(ring, seg, wH, a, 0, 0, segScale, invH, q4, tid);
#pragma unroll
        for (int ct = 0; ct < CTM; ct++) loadB(0, ct, 0);
        __syncthreads();
        // one step's products from ring buffer rb, pair half h, basis buffer buf; the basis
        // refills of step sn follow each column tile's products
        auto step_products = [&](int rb, int h, int buf, int sn, bool refill) {
            __builtin_amdgcn_sched_barrier(0);
            u16x8 ah[RTW][NT], oh[RTW][NT];
#pragma unroll
            for (int r = 0; r < RTW; r++) {
                const int slot = (rb * 2 * NT * GB + 2 * h + (lane >> 5)) * FT + (rt0 + r) * 32 + (lane & 31);
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    ah[r][t] = ring[slot + t * GB * FT];
                    oh[r][t] = ring[slot + (NT + t) * GB * FT];
                }
            }
            // the other buffer's refill goes out here, a whole step ahead of its use
            if (NB == 2 && refill) loadB(buf ^ 1, 0, sn);
            // keep the LDS reads (and that refill) ahead of the products (left alone, the
            // scheduler serialises each read with its first product and sinks the refill)
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int ct = 0; ct < CTM; ct++) {
                // X^T = basis (A: rows = bins) x frames (B: columns = frames); bin k from the
                // columns as stored, bin N/2 - k from the same columns with odd rows negated
#pragma unroll
                for (int r = 0; r < RTW; r++) mfma_split<PREC, NT>(acc[r][ct][0], bc[buf][ct], ah[r]);
#pragma unroll
                for (int r = 0; r < RTW; r++) mfma_split<PREC, NT>(acc[r][ct][1], bs[buf][ct], oh[r]);
                u16x8 mc[NT], ms[NT];
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    mc[t] = flip_odd(bc[buf][ct][t]);
                    ms[t] = flip_odd(bs[buf][ct][t]);
                }
#pragma unroll
                for (int r = 0; r < RTW; r++) mfma_split<PREC, NT>(acc[r][ct][2], ms, ah[r]);
#pragma unroll
                for (int r = 0; r < RTW; r++) mfma_split<PREC, NT>(acc[r][ct][3], mc, oh[r]);
                if (NB == 1 && refill) loadB(0, ct, sn);
            }
        };
        constexpr int UPT = 2 * GB * FT / 256;            // operand units per thread and step pair
        constexpr int MPS = (PREC == kPrecBF16 ? 4 : (PREC == kPrecBF16x6 ? 24 : 12)) * U;   // products per step and wave
        static_assert(kSPB == 2, "the loop below runs step pairs");
        // full pairs: straight-line, the next pair's operand build beside the products (unit i of
        // it loaded before step s0 + i and finished after it); the last pair's build and refills
        // past the end are harmless (zero rows into the idle buffer, a clamped refill)
        const int nPairs = nSteps / 2;
        for (int sb = 0; sb < nPairs; sb++) {
            const int rb = sb & 1;
            const int s0 = 2 * sb;
            PrepIn pin;
            prep_load<FT>(pin, seg, wH, a, s0 + 2, tid, invH);
            step_products(rb, 0, 0, s0 + 1, true);
            prep_finish<PREC, FT>(ring, pin, tid, rb ^ 1, segScale, q4);
            if (UPT == 2) prep_load<FT>(pin, seg, wH, a, s0 + 2, tid + 256, invH);
            interleave_valu<MPS>();
            step_products(rb, 1, NB == 2 ? 1 : 0, min(s0 + 2, nSteps - 1), true);
            if (UPT == 2) prep_finish<PREC, FT>(ring, pin, tid + 256, rb ^ 1, segScale, q4);
            interleave_valu<MPS>();
            __syncthreads();                            // next pair built; buffer rb free again
        }
        if (nSteps & 1) step_products(nPairs & 1, 0, 0, 0, false);   // odd last step
        q4Part[tid] = q4;
        __syncthreads();  // segment and operand ring are dead: the power planes may overwrite them

        // ---- power in lane, bin N/4 from the partial sums, stored as bf16 hi and lo planes
        //      [FT][pws2] for the mel product (each value split once)
        tid = opaque(threadIdx.x); wave = tid >> 6; lane = tid & 63;
        uint16_t *pwH = reinterpret_cast<uint16_t *>(lds);
        uint16_t *pwL = pwH + FT * pws2;
        // The planes hold the power of the scaled operands (segScale * basis scale, a power of two
        // for fp16x3, 1 otherwise); the mel output is multiplied by melScale = 1 / that scale (its
        // square for power 2).  Scaling by a power of two is exact, so this is the value the
        // unscaled power would give, with the multiply once per mel output instead of once per bin.
        const int sExp = (PREC == kPrecFP16x3) ? ilogbf(segScale) - ilogbf(a.basisInv) : 0;
        const float unscale = ldexpf(1.0f, -sExp);
        const float melScale = a.power == 1 ? unscale : unscale * unscale;
        {
            const int c0 = wave * U / NRT, r0 = wave * U % NRT;
            if (a.power == 1) store_power<true, RTW, CTM>(pwH, pwL, pws2, acc, c0, r0, a.N / 128, halfN, lane);
            else store_power<false, RTW, CTM>(pwH, pwL, pws2, acc, c0, r0, a.N / 128, halfN, lane);
        }
        const int Q = a.N / 4;
        if (tid < FT) {
            // bin N/4 from sums of the unscaled samples: brought to the scaled domain before squaring
            float re = 0.0f, im = 0.0f;
            for (int q = tid; q < NW * 64; q += FT) {
                re += q4Part[q].x;
                im += q4Part[q].y;
            }
            re = ldexpf(re, sExp);
            im = ldexpf(im, sExp);
            float p = 0.5f * (re * re + im * im);
            if (a.power == 1) p = sqrtf(p);
            put_power(pwH, pwL, tid * pws2 + Q, p);
        }
        {
            // positions the mel product reads that hold no bin: Q+1..Q+3 (a 2-byte and a 4-byte
            // store) and N/2+4 .. KP-1 (8-byte stores: N/2+4 and KP are multiples of 4), in both
            // planes; 16 threads per frame
            const int nc = (a.KP - (halfN + 4)) / 4;
            for (int e = tid; e < FT * 16; e += NW * 64) {
                const int row = (e >> 4) * pws2, c = e & 15;
                if (c < nc) {
                    *reinterpret_cast<uint2 *>(&pwH[row + halfN + 4 + 4 * c]) = make_uint2(0u, 0u);
                    *reinterpret_cast<uint2 *>(&pwL[row + halfN + 4 + 4 * c]) = make_uint2(0u, 0u);
                } else if (c == 15) {
                    pwH[row + Q + 1] = 0;
                    pwL[row + Q + 1] = 0;
                    *reinterpret_cast<uint32_t *>(&pwH[row + Q + 2]) = 0u;
                    *reinterpret_cast<uint32_t *>(&pwL[row + Q + 2]) = 0u;
                }
            }
        }
        __syncthreads();

        // ---- mel work list of this wave; the weight fragments of its first kWA entries are
        //      requested before the next tile's samples.  Slot u of the ring holds entry e with
        //      e % kWA == u; every refill is unconditional (an index past the end reloads the last
        //      entry, never used), so each slot stays in one set of registers and a product waits
        //      only for its own fragments.
        tid = opaque(threadIdx.x); wave = tid >> 6; lane = tid & 63;
        // the work list is the wave's own: its length and entries are wave-uniform (scalar loop
        // control and item-end branches)
        const int n = __builtin_amdgcn_readfirstlane(melSeqLen[wave]);
        const uint16_t *seq = melSeq[wave];
        u16x8 wh[kWA], wl[kWA];
        int ent[kWA];
        auto loadW = [&](int slot, int e) {
            const int en = __builtin_amdgcn_readfirstlane(seq[min(e, n - 1)]), mt = en >> 9, k0 = (en & 255) * 32;
            ent[slot] = en;
            const size_t o = (static_cast<size_t>(k0 / 8 + (lane >> 4)) * a.Mpad + mt * 16 + (lane & 15)) * 8;
            wh[slot] = *reinterpret_cast<const u16x8 *>(a.mHi + o);
            wl[slot] = *reinterpret_cast<const u16x8 *>(a.mLo + o);
        };
        // the next tile's record first: its load is then not queued behind the weight fragments
        const int nt = tile + gridDim.x;
        if (nt < T) cur = tile_info(a, nt, FT, span, lane);
        if (n > 0) {
#pragma unroll
            for (int u = 0; u < kWA; u++) loadW(u, u);
        }

        // ---- the next tile's samples load while the mel product runs (the accumulators are dead)
        tid = opaque(threadIdx.x); wave = tid >> 6; lane = tid & 63;
        if (nt < T) seg_prefetch(pf, cur, span, tid);
        else seg_prefetch_none(pf);

        // ---- mel^T = Wmel [M x KP] . pw^T [KP x FT] on MFMA 16x16x32 bf16 split hi+lo, over each
        tid = opaque(threadIdx.x); wave = tid >> 6; lane = tid & 63;
        //      16-filter tile's band of bins, FT/32 frame groups per work item
        if (n > 0) {
            constexpr int FGI = FT / 32;                     // frame groups per work item
            const bool fastLog = a.minRatio >= 1.17549435e-38f;   // max(minRatio, x) is never subnormal
            float melMax = -3.40282347e38f;
            f32x4 acc[FGI];
#pragma unroll
            for (int g = 0; g < FGI; g++) acc[g] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
            for (int e0 = 0; e0 < n; e0 += kWA) {
#pragma unroll
                for (int u = 0; u < kWA; u++) {
                    const int e = e0 + u;
                    if (e < n) {
                        const int item = ent[u] >> 8, mt = item >> 1, gb = (item & 1) * FGI, k0 = (ent[u] & 255) * 32;
                        const bf16x8 w0 = __builtin_bit_cast(bf16x8, wh[u]), w1 = __builtin_bit_cast(bf16x8, wl[u]);
                        const int q0 = (gb * 16 + (lane & 15)) * pws2 + k0 + 8 * (lane >> 4);
                        bf16x8 ph[FGI], pl[FGI];
#pragma unroll
                        for (int g = 0; g < FGI; g++) {
                            ph[g] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&pwH[q0 + g * 16 * pws2]));
                            pl[g] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8 *>(&pwL[q0 + g * 16 * pws2]));
                        }
#pragma unroll
                        for (int g = 0; g < FGI; g++) acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0, ph[g], acc[g], 0, 0, 0);
#pragma unroll
                        for (int g = 0; g < FGI; g++) acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0, pl[g], acc[g], 0, 0, 0);
#pragma unroll
                        for (int g = 0; g < FGI; g++) acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, ph[g], acc[g], 0, 0, 0);
                        // slot (u + 1) % kWA already holds entry e + 1
                        if (e + 1 == n || (ent[(u + 1) % kWA] >> 8) != item) {
                            // lane holds mel rows mt*16 + 4*(lane/16) + r of frame (gb+g)*16 + lane%16;
                            // the value to store is chosen once per item (uniform branches)
                            f32x4 val[FGI];
                            if (a.clipMax) {
#pragma unroll
                                for (int g = 0; g < FGI; g++) {
                                    val[g] = acc[g] * melScale;
                                    melMax = fmaxf(melMax, fmaxf(fmaxf(val[g][0], val[g][1]), fmaxf(val[g][2], val[g][3])));
                                }
                            } else if (fastLog) {
#pragma unroll
                                for (int g = 0; g < FGI; g++)
#pragma unroll
                                    for (int r = 0; r < 4; r++)
                                        val[g][r] = a.mult * __builtin_amdgcn_logf(fmaxf(a.minRatio, (acc[g][r] * melScale) * a.invRef));
                            } else {
#pragma unroll
                                for (int g = 0; g < FGI; g++)
#pragma unroll
                                    for (int r = 0; r < 4; r++) val[g][r] = db_fast(acc[g][r] * melScale, a);
                            }
                            const int m0 = mt * 16 + 4 * (lane >> 4);
                            const size_t rStride = a.ntf ? 1 : v.hs;   // between consecutive rows m
#pragma unroll
                            for (int g = 0; g < FGI; g++) {
                                const int fr = (gb + g) * 16 + (lane & 15);
                                if (f0 + fr < v.F) {
                                    float *o = out_at(a, v, m0, f0 + fr);
#pragma unroll
                                    for (int r = 0; r < 4; r++)
                                        if (m0 + r < a.M) o[r * rStride] = val[g][r];
                                }
                                acc[g] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                            }
                        }
                    }
                    loadW(u, e + kWA);
                }
            }
            if (a.clipMax) clip_max_update(a, v.b, melMax);
        }
        __syncthreads();  // the power planes are reused by the next tile's segment
    }
}

// ------------------------------------------------------------------------- dispatch
// Format of the folded basis a precision reads (DftTables::splitKind): none for FP32, fp16 scaled
// hi/lo for FP16X3, bf16 hi/mid/lo for BF16X6, bf16 hi/lo otherwise.
int split_kind(RpptLogMelPrecision precision)
{
    switch (precision) {
    case RPPT_LOGMEL_FP32: return -1;
    case RPPT_LOGMEL_FP16X3: return 1;
    case RPPT_LOGMEL_BF16X6: return 2;
    default: return 0;
    }
}

RppStatus check_params(const RpptLogMelParams *p, RpptLogMelPrecision precision)
{
    if (!p->multiplier) return RPP_ERROR_ZERO_DIVISION;
    const int nfft = p->nfft ? p->nfft : p->windowLength;
    if (p->windowLength <= 0 || p->windowStep <= 0 || nfft < 2 || p->windowLength > nfft || p->numFilter <= 0)
        return RPP_ERROR_INVALID_ARGUMENTS;
    if (p->maxFreq < 0 || p->maxFreq > p->sampleRate / 2 || p->minFreq < 0 || p->minFreq > p->sampleRate / 2)
        return RPP_ERROR_INVALID_ARGUMENTS;
    if (p->referenceMagnitude < 0.0f) return RPP_ERROR_INVALID_ARGUMENTS;
    if (precision < RPPT_LOGMEL_FP32 || precision > RPPT_LOGMEL_BF16X6) return RPP_ERROR_INVALID_ARGUMENTS;
    // the folded MFMA path: nfft 256, 512, 768 or 1024, window length a multiple of 16 (its half a
    // multiple of 8), hop a multiple of 8, at most 256 mel filters
    if (precision != RPPT_LOGMEL_FP32 &&
        ((p->windowStep % 8) != 0 || (p->windowLength % 16) != 0 || nfft > 1024 || (nfft % 256) != 0 ||
         p->numFilter > 256))
        return RPP_ERROR_NOT_IMPLEMENTED;
    return RPP_SUCCESS;
}

// Device memory of one launch supplied by the caller instead of the handle: the tables, and a
// workspace of `wsBytes` bytes (256-byte aligned) for the tile plan, tile records and clip maxima.
struct LaunchMem {
    const DftTables *dt;
    const MelTables *mt;
    char *ws;
    size_t wsBytes;
    const FftTables *ft;   // tables of the factored kernel (null: that kernel is not used)
    const MelFrag *mf;
};

std::string mel_key(const RpptLogMelParams *p, int K)
{
    char keyb[160];
    snprintf(keyb, sizeof(keyb), "%d/%a/%a/%a/%d/%d/%d", K, p->sampleRate, p->maxFreq, p->minFreq,
             static_cast<int>(p->melFormula), p->numFilter, p->normalize ? 1 : 0);
    return std::string(keyb);
}

// Validated parameters + buffers -> planner and fused kernel on h->stream.  Tables and workspace
// come from the handle (cached tables, grow-only workspace) unless `mem` supplies them; with `mem`
// the launch allocates nothing and fails with RPP_ERROR_OUT_OF_BOUND_SCRATCH_MEMORY_SIZE when its
// workspace does not fit.
RppStatus logmel_launch(rppHandle *h, const RpptLogMelParams *p, RpptLogMelPrecision precision, const float *src,
                        uint32_t srcStride, const long long *srcOff, const int *lengths, int B, float *dst,
                        uint32_t dstNStride, uint32_t hStride, const long long *dstOff, int ntf, uint32_t frameCap,
                        const int *begin = nullptr, const LaunchMem *mem = nullptr, const FrontEnd *fe = nullptr)
{
    const bool mfma = precision != RPPT_LOGMEL_FP32;
    const int nfft = p->nfft ? p->nfft : p->windowLength;
    const int K = nfft / 2 + 1;
    const int splitKind = split_kind(precision);
    const DftTables *dt = mem ? mem->dt : get_dft_tables(h, nfft, p->windowLength, nullptr, splitKind);
    const MelTables *mt = mem ? mem->mt
                              : get_mel_tables(h, K, p->sampleRate, p->maxFreq, p->minFreq,
                                               static_cast<int>(p->melFormula), p->numFilter, p->normalize ? 1 : 0);
    if (!dt || !mt) return RPP_ERROR;
    const bool clipRef = p->referenceMagnitude == 0.0f;
    // the factored-DFT kernel (logmel_fft.hip) where its launch shape and LDS budget allow
    const FftTables *ft = nullptr;
    const MelFrag *mf = nullptr;
    if (fft_eligible(p, precision, B) && !(h->tiles.blocksPerCU > 0 && h->tiles.framesPerTile > 0 &&
                                            h->tiles.framesPerTile != 16) &&
        (nfft == 512 || fft_r2_launchable(p, ntf, clipRef, fe != nullptr))) {
        ft = mem ? mem->ft : get_fft_tables(h, nfft, p->windowLength);
        mf = mem ? mem->mf : get_mel_frag(h, mt, mel_key(p, K));
        if (!ft || !mf || mf->nPairs <= 0 ||
            fft_lds_bytes(nfft, p->windowStep, B, mf->nPairs, begin != nullptr, srcOff != nullptr, nullptr,
                          fe != nullptr) > static_cast<size_t>(h->maxLdsOptin)) {
            ft = nullptr;
            mf = nullptr;
        }
    }
    // the fused front end needs the factored kernel's common launch; the caller falls back otherwise
    if (fe && (!ft || clipRef || ntf || p->power != 2 || begin || srcOff)) return RPP_ERROR_NOT_IMPLEMENTED;
    // tile shape: FP32 kernel 32 frames per tile; MFMA kernel 4 waves of CT = nfft/256 32-bin column
    // tiles each, 64 frames per tile up to nfft 512 and 32 above
    // (nfft 512 also runs 32-frame tiles, one unit per wave, when the handle's tiles ask for them)
    const int CT = nfft / 256;
    const int FT = mfma ? (CT == 2 && h->tiles.framesPerTile == 32 ? 32 : (CT <= 2 ? 64 : 32)) : kFT32;
    // per-tile records for the MFMA kernel on the strided layout, when their count is bounded by
    // the row stride and the frame cap (every clip has at most ceil(maxFrames / FT) tiles)
    // (records are only kept while they fit: 256 MB of the handle's workspace, or what a caller's
    // workspace leaves after the tile plan; otherwise the kernel searches the plan per tile)
    const size_t recOff = (sizeof(int) * (2 * B + 1) + 255) / 256 * 256;
    const size_t recCap = mem ? (mem->wsBytes > recOff ? mem->wsBytes - recOff : 0) : (static_cast<size_t>(256) << 20);
    size_t recBytes = 0;
    if (mfma && !srcOff) {
        const long long maxF = std::min<long long>(
            std::max(num_windows(static_cast<int>(std::min<uint32_t>(srcStride, 0x7fffffffu)), p->windowLength,
                                 p->windowStep, p->centerWindows), 0),
            frameCap);
        const size_t n = static_cast<size_t>(B) * static_cast<size_t>((maxF + FT - 1) / FT);
        if (n * sizeof(int4) <= recCap) recBytes = n * sizeof(int4);
    }
    // the MFMA kernel plans in each block (no planner launch, one dependent dispatch less) when the
    // per-tile records are bounded, the clip count fits the segment region and no per-clip maximum
    // needs initialising; each block's records take kPad slots, 256-byte aligned
    // (the three-term bf16 kernel keeps one block per CU: its operands take the whole register file)
    const int bpc = h->tiles.blocksPerCU > 0 ? h->tiles.blocksPerCU
                                             : (mfma && precision != RPPT_LOGMEL_BF16X6 ? 2 : 1);
    const unsigned grid = static_cast<unsigned>(h->numCU * bpc);
    bool selfPlan = recBytes > 0 && !clipRef && B <= 4096;
    int kPad = 0;
    size_t blockBytes = 0;
    if (selfPlan) {
        const size_t tiles = recBytes / sizeof(int4);
        kPad = static_cast<int>(((tiles + grid - 1) / grid + 15) / 16 * 16);
        blockBytes = static_cast<size_t>(grid) * kPad * sizeof(int4);
        if (blockBytes <= recCap) recBytes = 0;
        else { selfPlan = false; kPad = 0; blockBytes = 0; }
    }
    const size_t wsNeed = recOff + recBytes + blockBytes;
    char *ws = mem ? (wsNeed <= mem->wsBytes ? mem->ws : nullptr) : static_cast<char *>(workspace(h, wsNeed));
    if (!ws) return mem ? RPP_ERROR_OUT_OF_BOUND_SCRATCH_MEMORY_SIZE : RPP_ERROR_NOT_ENOUGH_MEMORY;
    int *tileStart = reinterpret_cast<int *>(ws);   // [B + 1], then the per-clip maxima [B]
    int4 *tileRec = recBytes ? reinterpret_cast<int4 *>(ws + recOff) : nullptr;
    int4 *blockRec = blockBytes ? reinterpret_cast<int4 *>(ws + recOff) : nullptr;

    LogMelArgs a{};
    a.src = src;
    a.srcStride = srcStride;
    a.lengths = lengths;
    a.begin = begin;
    a.B = B;
    a.dst = dst;
    a.dstNStride = dstNStride;
    a.hStride = hStride;
    a.srcOff = srcOff;
    a.dstOff = dstOff;
    a.coeff = p->preEmphasisCoeff;
    a.border = static_cast<int>(p->borderType);
    a.N = nfft;
    a.W = p->windowLength;
    a.H = p->windowStep;
    a.center = p->centerWindows;
    a.reflect = p->reflectPadding;
    a.power = p->power == 2 ? 2 : 1;   // rppt_spectrogram_gpu's rule: 2 squared magnitude, else magnitude
    a.K = K;
    a.M = p->numFilter;
    // dB constants exactly as cpu/kernel/to_decibels.cpp:38-43 computes them
    float minRatio = static_cast<float>(std::pow(10.0, static_cast<double>(p->cutOffDB / p->multiplier)));
    if (minRatio == 0.0f) minRatio = std::nextafter(0.0f, 1.0f);
    a.minRatio = minRatio;
    a.mult = p->multiplier * 0.3010299956639812f;
    a.invRef = clipRef ? 1.0f : 1.f / p->referenceMagnitude;
    a.ntf = ntf;
    a.frameCap = frameCap;
    a.clipMax = clipRef ? tileStart + B + 1 : nullptr;
    a.tileStart = tileStart;
    a.tileRec = tileRec;
    a.blockRec = blockRec;
    a.kPad = kPad;
    a.basis = dt->basis;
    a.cols = dt->cols;
    a.bHi = dt->foldHi;
    a.bLo = dt->foldLo;
    a.bLo2 = dt->foldLo2;
    a.wHalf = dt->wHalf;
    a.Wp = dt->Jp;
    a.basisInv = splitKind == 1 ? 1.0f / 4096.0f : 1.0f;
    a.weightsDown = mt->weightsDown;
    a.intervals = mt->intervals;
    a.normFactors = mt->normFactors;
    a.filterBins = mt->filterBins;
    a.normalize = p->normalize ? 1 : 0;
    a.mHi = mt->denseHi;
    a.mLo = mt->denseLo;
    a.KP = mt->KP;
    a.Mpad = mt->Mpad;

    if (!selfPlan && (!ft || clipRef))   // the factored kernel plans in LDS; the planner then only sets the clip maxima
        hipLaunchKernelGGL(plan_tiles_kernel, dim3(1), dim3(1024), 0, h->stream, lengths, begin, a.B, a.srcStride,
                       a.frameCap, a.W, a.H, a.center, FT, srcOff ? 1 : 0, tileStart, a.clipMax, tileRec);
    const int pws = a.KP + kSkew;
    if (ft) {
        FftLaunch L{};
        L.p = p;
        L.src = src;
        L.srcStride = srcStride;
        L.srcOff = srcOff;
        L.lengths = lengths;
        L.begin = begin;
        L.B = B;
        L.dst = dst;
        L.dstNStride = dstNStride;
        L.hStride = hStride;
        L.dstOff = dstOff;
        L.ntf = ntf;
        L.frameCap = frameCap;
        L.clipMax = a.clipMax;
        L.minRatio = a.minRatio;
        L.mult = a.mult;
        L.invRef = a.invRef;
        L.ft = ft;
        L.mf = mf;
        L.fe = fe;
        const RppStatus st = logmel_fft_launch(h, L);
        if (st != RPP_SUCCESS) return st;
    } else if (!mfma) {
        const int span = (kFT32 - 1) * a.H + a.W;
        const int segFloats = (seg_pos(span - 1, a.H) + 1 + 3) / 4 * 4;
        const int bFloats = kKC * kCC;
        const int pwFloats = kFT32 * pws;
        const int passes = (a.cols + kCC - 1) / kCC;
        const int alias = passes == 1 ? 1 : 0;
        const size_t bytes = sizeof(float) * (segFloats + (alias ? std::max(bFloats, pwFloats) : bFloats + pwFloats));
        if (bytes > static_cast<size_t>(h->maxLdsPerBlock) - kFT32 * kKC * sizeof(float))
            return RPP_ERROR_OUT_OF_BOUND_SHARED_MEMORY_SIZE;
        hipLaunchKernelGGL(logmel_fp32_kernel, dim3(grid), dim3(kThreads), bytes, h->stream, a, segFloats, bFloats,
                           pwFloats, alias);
    } else {
        const int span = (FT - 1) * a.H + a.W;
        const int segFloats = (seg_pos(span - 1, a.H) + 1 + 8 + 3) / 4 * 4;
        const size_t pwBytes = static_cast<size_t>(4) * FT * (a.KP + 8);   // bf16 hi + lo planes
        const size_t staticBytes = sizeof(float) * (4 + 2 * 4 * 64 + 4) + sizeof(uint16_t) * (4 * kMelSeq + 16);
        const size_t ringBytes = static_cast<size_t>(128) * split_terms(precision == RPPT_LOGMEL_BF16X6 ? kPrecBF16x6 : 0) *
                                 kSPB * FT;   // [2][2 NT][2 kSPB][FT] 16-byte slices
        const size_t alias = (std::max(sizeof(float) * segFloats + ringBytes, pwBytes) + 15) / 16 * 16;
        const int wOff = static_cast<int>(alias / sizeof(float));
        const size_t bytes = alias + sizeof(float) * ((a.W / 2 + 4 + 3) / 4 * 4);   // window half + 4 zeros
        if (bytes + staticBytes > static_cast<size_t>(h->maxLdsPerBlock)) return RPP_ERROR_OUT_OF_BOUND_SHARED_MEMORY_SIZE;
#define RPPA_LAUNCH_ONE(PR, F, C, R)                                                                            \
        do {                                                                                                    \
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&logmel_mfma_kernel<PR, F, C, R>),         \
                                      hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));     \
            hipLaunchKernelGGL((logmel_mfma_kernel<PR, F, C, R>), dim3(grid), dim3(kBlockMfma), bytes, h->stream, a, segFloats, wOff); \
        } while (0)
        // units (32-bin column tile of direct bins, 32-frame row tile) per wave: nfft 256: 2 x 2
        // over 4 waves; 512: 4 x 2 (or 4 x 1 at 32 frames); 768: 6 x 1 (two waves repeat a tile);
        // 1024: 8 x 1
#define RPPA_LAUNCH(PR)                                                                                         \
        do {                                                                                                    \
            if (CT == 1) RPPA_LAUNCH_ONE(PR, 64, 1, 1);                                                         \
            else if (CT == 2 && FT == 32) RPPA_LAUNCH_ONE(PR, 32, 1, 1);                                        \
            else if (CT == 2) RPPA_LAUNCH_ONE(PR, 64, 1, 2);                                                    \
            else RPPA_LAUNCH_ONE(PR, 32, 2, 1);                                                                 \
        } while (0)
        if (precision == RPPT_LOGMEL_BF16X3) RPPA_LAUNCH(kPrecBF16x3);
        else if (precision == RPPT_LOGMEL_FP16X3) RPPA_LAUNCH(kPrecFP16x3);
        else if (precision == RPPT_LOGMEL_BF16X6) RPPA_LAUNCH(kPrecBF16x6);
        else RPPA_LAUNCH(kPrecBF16);
#undef RPPA_LAUNCH
#undef RPPA_LAUNCH_ONE
    }
    RPPA_HIP(hipGetLastError());
    if (clipRef) {
        hipLaunchKernelGGL(db_from_clip_max_kernel, dim3(32, std::min(B, 65535)), dim3(256), 0, h->stream, a, mfma ? 0 : 1);
        RPPA_HIP(hipGetLastError());
    }
    return RPP_SUCCESS;
}

}  // namespace
}  // namespace rppa

using namespace rppa;

RppStatus check_logmel_params(const RpptLogMelParams *p, RpptLogMelPrecision precision)
{
    return check_params(p, precision);
}

RPPA_EXPORT RppStatus rppt_logmel_set_tiles(rppHandle_t h, const RpptLogMelTiles *t)
{
    if (!h || !t) return RPP_ERROR_INVALID_ARGUMENTS;
    if (t->framesPerTile < 0) return RPP_ERROR_INVALID_ARGUMENTS;
    if (t->blocksPerCU < 0 || t->blocksPerCU > 8) return RPP_ERROR_INVALID_ARGUMENTS;
    h->tiles = *t;
    return RPP_SUCCESS;
}

RPPA_EXPORT RppStatus rppt_logmel_get_tiles(rppHandle_t h, RpptLogMelTiles *t)
{
    if (!h || !t) return RPP_ERROR_INVALID_ARGUMENTS;
    *t = h->tiles;
    return RPP_SUCCESS;
}

namespace rppa {
namespace {
// Validation and launch shared by the two strided entry points (begin null: whole rows).
RppStatus fused_entry(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr, RpptDescPtr dstDescPtr,
                      const Rpp32s *srcBeginTensor, Rpp32s *srcLengthTensor, const RpptLogMelParams *p,
                      RpptLogMelPrecision precision, rppHandle_t h)
{
    if (!h || !srcDescPtr || !dstDescPtr || !p) return RPP_ERROR_INVALID_ARGUMENTS;
    if (srcDescPtr->numDims != 2) return RPP_ERROR_INVALID_SRC_DIMS;
    if (dstDescPtr->layout != NFT && dstDescPtr->layout != NTF) return RPP_ERROR_INVALID_DST_LAYOUT;
    if (srcDescPtr->dataType != F32 || dstDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcLengthTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    RppStatus st = check_params(p, precision);
    if (st != RPP_SUCCESS) return st;
    const unsigned B = srcDescPtr->n;
    if (B == 0) return RPP_SUCCESS;
    // NFT: rows are filters (h >= numFilter), a row holds up to hStride frames.
    // NTF: rows are frames (at most h of them are written), a row holds hStride >= numFilter values.
    const int ntf = dstDescPtr->layout == NTF ? 1 : 0;
    if (ntf ? dstDescPtr->strides.hStride < static_cast<unsigned>(p->numFilter)
            : dstDescPtr->h < static_cast<unsigned>(p->numFilter))
        return RPP_ERROR_INVALID_DST_DIMS;
    const uint32_t frameCap = ntf ? dstDescPtr->h : dstDescPtr->strides.hStride;
    return logmel_launch(h, p, precision, static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, nullptr,
                         srcLengthTensor, static_cast<int>(B), static_cast<float *>(dstPtr),
                         dstDescPtr->strides.nStride, dstDescPtr->strides.hStride, nullptr, ntf, frameCap,
                         srcBeginTensor);
}
}  // namespace
}  // namespace rppa

RPPA_EXPORT RppStatus rppt_logmel_fused_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                            RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor,
                                            const RpptLogMelParams *p, RpptLogMelPrecision precision,
                                            rppHandle_t h)
{
    return rppa::fused_entry(srcPtr, srcDescPtr, dstPtr, dstDescPtr, nullptr, srcLengthTensor, p, precision, h);
}

// The fused launch over a region of each clip's row: samples [srcBeginTensor[b], + srcLengthTensor[b])
// (begin clamped to the row, the length to what remains of it), e.g. the begin/length pair
// rppt_non_silent_region_detection_gpu writes.  Same output as rppt_logmel_fused_gpu on the
// trimmed clips, with no copy of them.
RPPA_EXPORT RppStatus rppt_logmel_fused_region_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                                   RpptDescPtr dstDescPtr, Rpp32s *srcBeginTensor,
                                                   Rpp32s *srcLengthTensor, const RpptLogMelParams *p,
                                                   RpptLogMelPrecision precision, rppHandle_t h)
{
    if (!srcBeginTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    return rppa::fused_entry(srcPtr, srcDescPtr, dstPtr, dstDescPtr, srcBeginTensor, srcLengthTensor, p, precision, h);
}

// Down-mix, resample and the fused log-mel: interleaved clips at inRate in, dB at outRate out -- the
// chain rppt_down_mix_resample_gpu -> rppt_logmel_fused_gpu over the resampled clips.  With oneLaunch
// and where the factored-DFT kernel's common launch applies (fp16x3, nfft 512, hop 160, power 2,
// NFT, a fixed reference), inRate / outRate is 2 or 3 and the window has 16 lobes, it is one launch
// whose work items compute the down-mixed, resampled samples they read; otherwise the two launches
// through a buffer of the handle.  Both forms give the same bits.
RPPA_EXPORT RppStatus rppt_logmel_front_end_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                                RpptDescPtr dstDescPtr, Rpp32s *srcDimsTensor, Rpp32f inRate,
                                                Rpp32f outRate, bool normalizeWeights, RpptResamplingWindow &window,
                                                const RpptLogMelParams *p, RpptLogMelPrecision precision,
                                                Rpp32s oneLaunch, rppHandle_t h)
{
    if (!h || !srcDescPtr || !dstDescPtr || !p) return RPP_ERROR_INVALID_ARGUMENTS;
    const size_t dims = srcDescPtr->numDims - 1;
    if (dims != 1 && dims != 2) return RPP_ERROR_INVALID_SRC_DIMS;
    if (dstDescPtr->layout != NFT && dstDescPtr->layout != NTF) return RPP_ERROR_INVALID_DST_LAYOUT;
    if (srcDescPtr->dataType != F32 || dstDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcDimsTensor || !window.lookupPinned) return RPP_ERROR_INVALID_ARGUMENTS;
    if (window.lookupSize < 2 || window.lookupSize > 16384 || window.lobes < 0) return RPP_ERROR_INVALID_ARGUMENTS;
    if (!(inRate > 0.0f) || !(outRate > 0.0f)) return RPP_ERROR_INVALID_ARGUMENTS;
    RppStatus st = check_params(p, precision);
    if (st != RPP_SUCCESS) return st;
    const unsigned B = srcDescPtr->n;
    if (B == 0) return RPP_SUCCESS;
    const int ntf = dstDescPtr->layout == NTF ? 1 : 0;
    if (ntf ? dstDescPtr->strides.hStride < static_cast<unsigned>(p->numFilter)
            : dstDescPtr->h < static_cast<unsigned>(p->numFilter))
        return RPP_ERROR_INVALID_DST_DIMS;
    const uint32_t frameCap = ntf ? dstDescPtr->h : dstDescPtr->strides.hStride;
    const float *src = static_cast<const float *>(srcPtr);
    const uint32_t srcStride = srcDescPtr->strides.nStride;
    float *dst = static_cast<float *>(dstPtr);
    const double ratio = static_cast<double>(inRate) / outRate;
    const int k = static_cast<int>(ratio);
    if (oneLaunch && static_cast<double>(k) == ratio && fft_front_eligible(k, window.lobes)) {
        const FrontEnd fe{src, srcStride, srcDimsTensor, inRate, outRate, k, normalizeWeights ? 1 : 0,
                          window.lookupPinned, window.scale, window.center, window.lookupSize};
        st = logmel_launch(h, p, precision, src, srcStride, nullptr, nullptr, static_cast<int>(B), dst,
                           dstDescPtr->strides.nStride, dstDescPtr->strides.hStride, nullptr, ntf, frameCap, nullptr,
                           nullptr, &fe);
        if (st != RPP_ERROR_NOT_IMPLEMENTED) return st;
    }
    // two launches: every clip has at most srcStride samples per channel, so its resampled length
    // ceil(L * outRate / inRate) fits a row of rs
    const uint32_t rs = static_cast<uint32_t>(std::ceil(static_cast<double>(srcStride) * outRate / inRate)) + 8;
    const size_t lenOff = (static_cast<size_t>(B) * rs * sizeof(float) + 255) / 256 * 256;
    char *buf = static_cast<char *>(front_workspace(h, lenOff + sizeof(int) * B));
    if (!buf) return RPP_ERROR_NOT_ENOUGH_MEMORY;
    float *res = reinterpret_cast<float *>(buf);
    int *lens = reinterpret_cast<int *>(buf + lenOff);
    RPPA_HIP(down_mix_resample_launch(h->stream, src, srcStride, res, rs, B, srcDimsTensor, inRate, outRate, window,
                                      normalizeWeights ? 1 : 0, lens));
    return logmel_launch(h, p, precision, res, rs, nullptr, lens, static_cast<int>(B), dst, dstDescPtr->strides.nStride,
                         dstDescPtr->strides.hStride, nullptr, ntf, frameCap);
}

// Autotuner: times every tile shape the launch supports on the caller's batch (events on the
// handle's stream, `iters` launches each after one warm-up) and keeps the fastest in the handle.
RPPA_EXPORT RppStatus rppt_logmel_autotune(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                           RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor,
                                           const RpptLogMelParams *p, RpptLogMelPrecision precision, Rpp32s iters,
                                           rppHandle_t h, RpptLogMelTiles *chosen, Rpp32f *bestMs)
{
    if (!h || !p || iters <= 0) return RPP_ERROR_INVALID_ARGUMENTS;
    const int nfft = p->nfft ? p->nfft : p->windowLength;
    std::vector<RpptLogMelTiles> cands;
    // three axes: the kernel (the factored DFT, 16 frames per wave item and one 8-wave workgroup
    // per CU, where it applies; else the dense kernel), the dense kernel's frames per tile (64 or
    // 32 at nfft 512, which also sets how many (column tile, row tile) units each wave holds) and
    // its persistent workgroups per CU (1-4)
    if (srcDescPtr && fft_eligible(p, precision, static_cast<int>(srcDescPtr->n))) cands.push_back({16, 1});
    std::vector<int> fts;
    if (precision == RPPT_LOGMEL_FP32) fts = {kFT32};
    else if (nfft >= 768) fts = {32};
    else if (nfft / 256 == 2) fts = {64, 32};
    else fts = {64};
    for (int ft : fts)
        for (int bpc = 1; bpc <= 4; bpc++) cands.push_back({ft, bpc});
    hipEvent_t e0, e1;
    RPPA_HIP(hipEventCreate(&e0));
    RPPA_HIP(hipEventCreate(&e1));
    const RpptLogMelTiles saved = h->tiles;
    float best = 1e30f;
    RpptLogMelTiles bestT = saved;
    for (const RpptLogMelTiles &c : cands) {
        h->tiles = c;
        RppStatus st = rppt_logmel_fused_gpu(srcPtr, srcDescPtr, dstPtr, dstDescPtr, srcLengthTensor, p, precision, h);
        if (st != RPP_SUCCESS) { h->tiles = saved; return st; }
        RPPA_HIP(hipEventRecord(e0, h->stream));
        for (int i = 0; i < iters; i++)
            (void)rppt_logmel_fused_gpu(srcPtr, srcDescPtr, dstPtr, dstDescPtr, srcLengthTensor, p, precision, h);
        RPPA_HIP(hipEventRecord(e1, h->stream));
        RPPA_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        RPPA_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= iters;
        if (ms < best) { best = ms; bestT = c; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    h->tiles = bestT;
    if (chosen) *chosen = bestT;
    if (bestMs) *bestMs = best;
    return RPP_SUCCESS;
}

RPPA_EXPORT RppStatus rppt_logmel_fused_packed_gpu(const Rpp32f *srcPtr, const long long *srcOffsets,
                                                   const Rpp32s *srcLengthTensor, Rpp32s batchSize, Rpp32f *dstPtr,
                                                   const long long *dstOffsets, const RpptLogMelParams *p,
                                                   RpptLogMelPrecision precision, rppHandle_t h)
{
    if (!h || !p || !srcPtr || !srcOffsets || !srcLengthTensor || !dstPtr || !dstOffsets || batchSize < 0)
        return RPP_ERROR_INVALID_ARGUMENTS;
    RppStatus st = check_params(p, precision);
    if (st != RPP_SUCCESS) return st;
    if (batchSize == 0) return RPP_SUCCESS;
    return logmel_launch(h, p, precision, srcPtr, 0, srcOffsets, srcLengthTensor, batchSize, dstPtr, 0, 0,
                         dstOffsets, 0, 0);
}

// ------------------------------------------------------------------------- no-allocation entry
// rppt_logmel_fused_noalloc_gpu: the launch of rppt_logmel_fused_gpu with every byte of device
// memory beyond the caller's inputs and outputs -- the DFT and mel tables, the tile plan, the tile
// records and the per-clip maxima -- in one file-scope device array, so a call never allocates.
// Each device holds its own copy of the array until the process ends; host code reaches it only
// through the address hipGetSymbolAddress returns for the current device.
namespace rppa {
namespace {

constexpr size_t kScratchBytes = static_cast<size_t>(16) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ScratchDevice {
    bool resolved = false;
    char *ptr = nullptr;          // device address of `scratch` on this device
    std::string tableKey;         // parameter set whose tables the array holds ("" none)
    DftTables dt;
    MelTables mt;
    FftTables ft;
    MelFrag mf;
    bool hasFft = false;          // the factored kernel's tables are there too
    size_t tableBytes = 0;        // bytes at the start of the array taken by those tables
};
std::mutex gScratchMu;
std::map<int, ScratchDevice> gScratch;

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// Carves the tables of one parameter set out of the array's head (256-byte aligned pieces);
// returns the bytes used, or 0 when they do not fit.
size_t layout_tables(char *base, int nfft, int W, int splitKind, bool wantBasis, int numBins, int M, float **win,
                     int **range, DftTables &dt, MelTables &mt)
{
    size_t o = 0;
    auto take = [&](size_t bytes) {
        char *q = base + o;
        o += align256(bytes);
        return q;
    };
    *win = reinterpret_cast<float *>(take(sizeof(float) * W));
    if (wantBasis) dt.basis = reinterpret_cast<float *>(take(sizeof(float) * static_cast<size_t>(W) * (nfft + (nfft & 1))));
    if (splitKind >= 0) {
        const size_t n = dft_fold_elems(nfft, W);
        dt.foldHi = reinterpret_cast<uint16_t *>(take(n * 2));
        dt.foldLo = reinterpret_cast<uint16_t *>(take(n * 2));
        if (splitKind == 2) dt.foldLo2 = reinterpret_cast<uint16_t *>(take(n * 2));
        dt.wHalf = reinterpret_cast<float *>(take(sizeof(float) * (W / 2)));
    }
    mt.weightsDown = reinterpret_cast<float *>(take(sizeof(float) * numBins));
    mt.intervals = reinterpret_cast<int *>(take(sizeof(int) * numBins));
    mt.normFactors = reinterpret_cast<float *>(take(sizeof(float) * M));
    mt.filterBins = reinterpret_cast<int *>(take(sizeof(int) * 2 * M));
    mt.dense = reinterpret_cast<float *>(take(sizeof(float) * static_cast<size_t>(M) * numBins));
    const size_t ns = mel_dense_elems(numBins, M);
    mt.denseHi = reinterpret_cast<uint16_t *>(take(ns * 2));
    mt.denseLo = reinterpret_cast<uint16_t *>(take(ns * 2));
    *range = reinterpret_cast<int *>(take(sizeof(int) * 2));
    return o <= kScratchBytes ? o : 0;
}

// The factored kernel's tables after the others (from byte `o`); returns the new end, 0 if past the array.
size_t layout_fft_tables(char *base, size_t o, int nfft, int W, int M, FftTables &ft, MelFrag &mf)
{
    fft_tables_carve(base + o, W, ft, nfft);
    o += align256(fft_table_bytes(W, nfft));
    mf.frag = reinterpret_cast<uint16_t *>(base + o);
    o += align256(mel_frag_bytes(M, nfft));
    mf.pairOff = reinterpret_cast<int *>(base + o);
    mf.w256 = nfft == 1024 ? reinterpret_cast<float *>(mf.pairOff + 257) : nullptr;
    o += align256(mel_frag_word_bytes(nfft));
    return o <= kScratchBytes ? o : 0;
}

}  // namespace
}  // namespace rppa

RPPA_EXPORT RppStatus rppt_logmel_fused_noalloc_gpu(RppPtr_t srcPtr, RpptDescPtr srcDescPtr, RppPtr_t dstPtr,
                                                    RpptDescPtr dstDescPtr, Rpp32s *srcLengthTensor,
                                                    const RpptLogMelParams *p, RpptLogMelPrecision precision,
                                                    rppHandle_t h)
{
    using namespace rppa;
    if (!h || !srcDescPtr || !dstDescPtr || !p) return RPP_ERROR_INVALID_ARGUMENTS;
    if (h->backend != RPP_HIP_BACKEND) return RPP_ERROR_NOT_IMPLEMENTED;
    if (srcDescPtr->numDims != 2) return RPP_ERROR_INVALID_SRC_DIMS;
    if (dstDescPtr->layout != NFT && dstDescPtr->layout != NTF) return RPP_ERROR_INVALID_DST_LAYOUT;
    if (srcDescPtr->dataType != F32 || dstDescPtr->dataType != F32) return RPP_ERROR_NOT_IMPLEMENTED;
    if (!srcPtr || !dstPtr || !srcLengthTensor) return RPP_ERROR_INVALID_ARGUMENTS;
    RppStatus st = check_params(p, precision);
    if (st != RPP_SUCCESS) return st;
    const unsigned B = srcDescPtr->n;
    if (B == 0) return RPP_SUCCESS;
    const int ntf = dstDescPtr->layout == NTF ? 1 : 0;
    if (ntf ? dstDescPtr->strides.hStride < static_cast<unsigned>(p->numFilter)
            : dstDescPtr->h < static_cast<unsigned>(p->numFilter))
        return RPP_ERROR_INVALID_DST_DIMS;
    const uint32_t frameCap = ntf ? dstDescPtr->h : dstDescPtr->strides.hStride;

    int dev = 0;
    RPPA_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(gScratchMu);
    ScratchDevice &sd = gScratch[dev];
    if (!sd.resolved) {
        // once per device, with that device current: the array's device address (the host-side
        // name `scratch` is not one)
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            fprintf(stderr, "rpp_audio: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return RPP_ERROR;
        }
        sd.ptr = static_cast<char *>(ptr);
        sd.resolved = true;
    }

    // the tables of this parameter set, rebuilt in place when the array holds another set's
    const bool mfma = precision != RPPT_LOGMEL_FP32;
    const int nfft = p->nfft ? p->nfft : p->windowLength;
    const int K = nfft / 2 + 1;
    const int splitKind = split_kind(precision);
    if (splitKind >= 0 && ((nfft & 1) || (p->windowLength & 1))) return RPP_ERROR_NOT_IMPLEMENTED;
    char keyb[200];
    snprintf(keyb, sizeof(keyb), "%d/%d/%d/%d/%a/%a/%a/%d/%d/%d", nfft, p->windowLength, splitKind, K, p->sampleRate,
             p->maxFreq, p->minFreq, static_cast<int>(p->melFormula), p->numFilter, p->normalize ? 1 : 0);
    if (sd.tableKey != keyb) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(h->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
            fprintf(stderr, "rpp_audio: rppt_logmel_fused_noalloc_gpu must run once with these parameters before "
                            "stream capture (its tables are built then)\n");
            return RPP_ERROR;
        }
        DftTables dt;
        MelTables mt;
        float *win = nullptr;
        int *range = nullptr;
        const size_t used = layout_tables(sd.ptr, nfft, p->windowLength, splitKind, !mfma, K, p->numFilter, &win,
                                          &range, dt, mt);
        if (!used) return RPP_ERROR_OUT_OF_BOUND_SCRATCH_MEMORY_SIZE;
        sd.tableKey.clear();
        if (build_dft_tables_at(h->stream, nfft, p->windowLength, nullptr, splitKind, !mfma, win, dt) != hipSuccess ||
            build_mel_tables_at(h->stream, K, p->sampleRate, p->maxFreq, p->minFreq, static_cast<int>(p->melFormula),
                                p->numFilter, p->normalize ? 1 : 0, range, mt) != hipSuccess)
            return RPP_ERROR;
        // the factored kernel's tables too, when its launch shape fits these parameters (its mel pair
        // count is read back once, here, outside any capture)
        FftTables ft;
        MelFrag mf;
        bool hasFft = false;
        size_t end = used;
        if (fft_eligible(p, precision, 1)) {
            const size_t e2 = layout_fft_tables(sd.ptr, used, nfft, p->windowLength, p->numFilter, ft, mf);
            if (e2) {
                int n[2] = {0, 0};   // [255] tiles of a chunk not contiguous, [256] pair count
                if (build_fft_tables_at(h->stream, p->windowLength, ft) != hipSuccess ||
                    build_mel_frag_at(h->stream, mt, mf) != hipSuccess ||
                    hipMemcpyAsync(n, mf.pairOff + 255, sizeof(n), hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
                    hipStreamSynchronize(h->stream) != hipSuccess)
                    return RPP_ERROR;
                mf.nPairs = n[0] ? 0 : n[1];
                hasFft = true;
                end = e2;
            }
        }
        sd.dt = dt;
        sd.mt = mt;
        sd.ft = ft;
        sd.mf = mf;
        sd.hasFft = hasFft;
        sd.tableBytes = end;
        sd.tableKey = keyb;
    }
    LaunchMem mem{&sd.dt, &sd.mt, sd.ptr + sd.tableBytes, kScratchBytes - sd.tableBytes,
                  sd.hasFft ? &sd.ft : nullptr, sd.hasFft ? &sd.mf : nullptr};
    return logmel_launch(h, p, precision, static_cast<const float *>(srcPtr), srcDescPtr->strides.nStride, nullptr,
                         srcLengthTensor, static_cast<int>(B), static_cast<float *>(dstPtr),
                         dstDescPtr->strides.nStride, dstDescPtr->strides.hStride, nullptr, ntf, frameCap, nullptr,
                         &mem);
}
